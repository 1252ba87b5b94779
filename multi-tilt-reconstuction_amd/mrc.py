"""MRC / TXBR files through the library's file boundary (include/mtr_io.h)."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from ._lib import MtrError, VolumeSize, lib

HEADER_DTYPE = np.dtype([
    ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("mode", "<i4"),
    ("nxstart", "<i4"), ("nystart", "<i4"), ("nzstart", "<i4"),
    ("mx", "<i4"), ("my", "<i4"), ("mz", "<i4"),
    ("xlen", "<f4"), ("ylen", "<f4"), ("zlen", "<f4"),
    ("alpha", "<f4"), ("beta", "<f4"), ("gamma", "<f4"),
    ("mapc", "<i4"), ("mapr", "<i4"), ("maps", "<i4"),
    ("amin", "<f4"), ("amax", "<f4"), ("amean", "<f4"),
    ("ispg", "<i2"), ("nsymbt", "<i2"), ("next", "<i4"), ("creatid", "<i2"), ("blank", "V30"),
    ("nint", "<i2"), ("nreal", "<i2"), ("sub", "<i2"), ("zfac", "<i2"),
    ("min2", "<f4"), ("max2", "<f4"), ("min3", "<f4"), ("max3", "<f4"), ("min4", "<f4"), ("max4", "<f4"),
    ("idtype", "<i2"), ("lens", "<i2"), ("nd1", "<i2"), ("nd2", "<i2"), ("vd1", "<i2"), ("vd2", "<i2"),
    ("tiltangles", "<f4", (6,)), ("xorg", "<f4"), ("yorg", "<f4"), ("zorg", "<f4"),
    ("cmap", "S4"), ("stamp", "V4"), ("rms", "<f4"), ("nlabl", "<i4"), ("labels", "S80", (10,)),
])
assert HEADER_DTYPE.itemsize == 1024

EMPTY = 999999


def _chk(rc, what):
    if rc != 0:
        raise MtrError(rc, f"{what}: {lib.mtr_last_error().decode(errors='replace')}")


def read_header(path: str) -> np.void:
    buf = np.zeros(1, HEADER_DTYPE)
    _chk(lib.mtr_mrc_read_header(os.fsencode(path), buf.ctypes.data_as(ct.c_void_p)), "mrc header")
    return buf[0]


def read(path: str) -> np.ndarray:
    h = read_header(path)
    out = np.empty((int(h["nz"]), int(h["ny"]), int(h["nx"])), np.float32)
    _chk(lib.mtr_mrc_read_data(os.fsencode(path), out.ctypes.data_as(ct.POINTER(ct.c_float)), out.size), "mrc data")
    return out


def write_volume(path: str, vol: VolumeSize, data) -> None:
    d = np.ascontiguousarray(data, np.float32)
    _chk(lib.mtr_mrc_write_volume(os.fsencode(path), ct.byref(vol), d.ctypes.data_as(ct.POINTER(ct.c_float))),
         "mrc write")


def write_stack(path: str, data) -> None:
    d = np.ascontiguousarray(data, np.float32)
    nz, ny, nx = d.shape
    _chk(lib.mtr_mrc_write_stack(os.fsencode(path), nx, ny, nz, d.ctypes.data_as(ct.POINTER(ct.c_float))),
         "mrc write")


def read_txbr(path: str, max_views: int = 4096, vol: VolumeSize | None = None):
    v = vol if vol is not None else VolumeSize(EMPTY, -EMPTY, EMPTY, -EMPTY, EMPTY, -EMPTY, 0, 0, 0)
    xc = np.zeros(4 * max_views)
    yc = np.zeros(4 * max_views)
    n = ct.c_int()
    _chk(lib.mtr_txbr_read(os.fsencode(path), ct.byref(v), xc.ctypes.data_as(ct.POINTER(ct.c_double)),
                           yc.ctypes.data_as(ct.POINTER(ct.c_double)), max_views, ct.byref(n)), "txbr read")
    return v, xc[: 4 * n.value].copy(), yc[: 4 * n.value].copy()


def write_txbr(path: str, vol: VolumeSize, xc, yc) -> None:
    xc = np.ascontiguousarray(xc, np.float64)
    yc = np.ascontiguousarray(yc, np.float64)
    _chk(lib.mtr_txbr_write(os.fsencode(path), ct.byref(vol), xc.ctypes.data_as(ct.POINTER(ct.c_double)),
                            yc.ctypes.data_as(ct.POINTER(ct.c_double)), xc.size // 4), "txbr write")
