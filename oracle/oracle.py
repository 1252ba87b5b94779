"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, as the checker.
The product library never loads it.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrVol(ct.Structure):
    _fields_ = [(n, ct.c_int) for n in ("Xstart", "Xend", "Ystart", "Yend", "Zstart", "Zend", "X", "Y", "Z")]


class OrPrj(ct.Structure):
    _fields_ = [("X", ct.c_int), ("Y", ct.c_int), ("AngN", ct.c_int)]


class OrSlice(ct.Structure):
    _fields_ = [("z_start", ct.c_int), ("z_end", ct.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ct.CDLL(LIB_PATH)
    pf, pd, pi = ct.POINTER(ct.c_float), ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)
    pv, pp = ct.POINTER(OrVol), ct.POINTER(OrPrj)
    ppf, ppd = ct.POINTER(pf), ct.POINTER(pd)
    protos = {
        "or_weight": (None, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, pd, pd, pi, pi, pd, pd]),
        "or_divide_slice": (OrSlice, [ct.c_int] * 4),
        "or_fp": (ct.c_int, [pp, pv, pd, pd, pf, pf, pf]),
        "or_fp_slab": (ct.c_int, [pp, pv, ct.c_int, ct.c_int, pd, pd, pf, pf, pf]),
        "or_residual": (ct.c_int, [ct.c_int64, pf, pf, pf, pf]),
        "or_bp": (ct.c_int, [pp, pv, pd, pd, pf, pf, ct.c_float]),
        "or_bpt": (ct.c_int, [pp, pv, pd, pd, pf, pf]),
        "or_sirt": (ct.c_int, [pp, pv, pd, pd, pf, pf, ct.c_int, ct.c_float]),
        "or_form_z": (ct.c_int, [ct.c_int64, pf, pf, pf]),
        "or_mann": (ct.c_int, [ct.c_int64, pf, pf, pf, ct.c_float]),
        "or_mean": (ct.c_int, [ct.c_int64, ct.c_int, ppf, pf]),
        "or_mace": (ct.c_int, [ct.c_int, pp, pv, ppd, ppd, ppf, ct.c_int, ct.c_int, ct.c_float, ct.c_float, pf]),
        "or_fp_f64": (ct.c_int, [pp, pv, pd, pd, pd, pd, pd]),
        "or_bp_sums_volume_f64": (ct.c_int, [pp, pv, pd, pd, pd, pd, pd]),
        "or_bp_f64": (ct.c_int, [pp, pv, pd, pd, pd, pd, ct.c_double]),
        "or_bpt_f64": (ct.c_int, [pp, pv, pd, pd, pf, pd]),
        "or_sirt_f64": (ct.c_int, [pp, pv, pd, pd, pf, pd, ct.c_int, ct.c_double]),
        "or_mace_f64": (ct.c_int, [ct.c_int, pp, pv, ppd, ppd, ppf, ct.c_int, ct.c_int, ct.c_double, ct.c_double, pd]),
        "or_set_threads": (ct.c_int, [ct.c_int]),
    }
    for n, (r, a) in protos.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _f(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def _d(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def vol_of(v) -> OrVol:
    return OrVol(v.Xstart, v.Xend, v.Ystart, v.Yend, v.Zstart, v.Zend, v.X, v.Y, v.Z)


def prj_of(p) -> OrPrj:
    return OrPrj(p.X, p.Y, p.AngN)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> int:
    return lib.or_set_threads(n)


def weight(X, Y, Z, a, xc, yc):
    i0, k0, du, dv = ct.c_int(), ct.c_int(), ct.c_double(), ct.c_double()
    xc, yc = _c64(xc), _c64(yc)
    lib.or_weight(X, Y, Z, a, _d(xc), _d(yc), ct.byref(i0), ct.byref(k0), ct.byref(du), ct.byref(dv))
    return i0.value, k0.value, du.value, dv.value


def divide_slice(myid, n, Z, S):
    s = lib.or_divide_slice(myid, n, Z, S)
    return s.z_start, s.z_end


def fp(prj, vol, xc, yc, x):
    x = _c32(x)
    s = np.empty((prj.AngN, prj.Y, prj.X), np.float32)
    c = np.empty_like(s)
    xc, yc = _c64(xc), _c64(yc)
    rc = lib.or_fp(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(x), _f(s), _f(c))
    assert rc == 0
    return s, c


def fp_slab(prj, vol, zs, ze, xc, yc, x):
    x = _c32(x)
    s = np.empty((prj.AngN, prj.Y, prj.X), np.float32)
    c = np.empty_like(s)
    xc, yc = _c64(xc), _c64(yc)
    rc = lib.or_fp_slab(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), zs, ze, _d(xc), _d(yc), _f(x), _f(s), _f(c))
    assert rc == 0
    return s, c


def residual(p, s, c):
    p, s, c = _c32(p), _c32(s), _c32(c)
    r = np.empty_like(p)
    lib.or_residual(p.size, _f(p), _f(s), _f(c), _f(r))
    return r


def bp(prj, vol, xc, yc, r, x, lam):
    r = _c32(r)
    x = _c32(x).copy()
    xc, yc = _c64(xc), _c64(yc)
    lib.or_bp(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(r), _f(x), lam)
    return x


def bpt(prj, vol, xc, yc, p):
    p = _c32(p)
    x = np.empty((vol.Z, vol.Y, vol.X), np.float32)
    xc, yc = _c64(xc), _c64(yc)
    lib.or_bpt(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(p), _f(x))
    return x


def sirt(prj, vol, xc, yc, p, x0, iters, lam):
    p = _c32(p)
    x = _c32(x0).copy()
    xc, yc = _c64(xc), _c64(yc)
    rc = lib.or_sirt(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(p), _f(x), iters, lam)
    assert rc == 0
    return x


def form_z(w, wbar):
    w, wbar = _c32(w), _c32(wbar)
    z = np.empty_like(w)
    lib.or_form_z(w.size, _f(w), _f(wbar), _f(z))
    return z


def mann(v, w, z, rho):
    v, z = _c32(v), _c32(z)
    w = _c32(w).copy()
    lib.or_mann(w.size, _f(v), _f(w), _f(z), rho)
    return w


def mean(ws):
    ws = [_c32(w) for w in ws]
    arr = (ct.POINTER(ct.c_float) * len(ws))(*[_f(w) for w in ws])
    out = np.empty_like(ws[0])
    lib.or_mean(ws[0].size, len(ws), arr, _f(out))
    return out


def mace(prjs, vol, xcs, ycs, ps, M, S, lam, rho):
    K = len(prjs)
    xcs = [_c64(a) for a in xcs]
    ycs = [_c64(a) for a in ycs]
    ps = [_c32(a) for a in ps]
    parr = (OrPrj * K)(*[prj_of(p) for p in prjs])
    xa = (ct.POINTER(ct.c_double) * K)(*[_d(a) for a in xcs])
    ya = (ct.POINTER(ct.c_double) * K)(*[_d(a) for a in ycs])
    pa = (ct.POINTER(ct.c_float) * K)(*[_f(a) for a in ps])
    out = np.empty((vol.Z, vol.Y, vol.X), np.float32)
    rc = lib.or_mace(K, parr, ct.byref(vol_of(vol)), xa, ya, pa, M, S, lam, rho, _f(out))
    assert rc == 0
    return out


# ---- f64 mode --------------------------------------------------------------
def fp_f64(prj, vol, xc, yc, x):
    x = _c64(x)
    s = np.empty((prj.AngN, prj.Y, prj.X), np.float64)
    c = np.empty_like(s)
    xc, yc = _c64(xc), _c64(yc)
    lib.or_fp_f64(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _d(x), _d(s), _d(c))
    return s, c


def bp_sums_f64(prj, vol, xc, yc, r):
    r = _c64(r)
    S = np.empty((vol.Z, vol.Y, vol.X), np.float64)
    C = np.empty_like(S)
    xc, yc = _c64(xc), _c64(yc)
    lib.or_bp_sums_volume_f64(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _d(r), _d(S), _d(C))
    return S, C


def bp_f64(prj, vol, xc, yc, r, x, lam):
    r = _c64(r)
    x = _c64(x).copy()
    xc, yc = _c64(xc), _c64(yc)
    lib.or_bp_f64(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _d(r), _d(x), lam)
    return x


def bpt_f64(prj, vol, xc, yc, p):
    p = _c32(p)
    x = np.empty((vol.Z, vol.Y, vol.X), np.float64)
    xc, yc = _c64(xc), _c64(yc)
    lib.or_bpt_f64(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(p), _d(x))
    return x


def sirt_f64(prj, vol, xc, yc, p, x0, iters, lam):
    p = _c32(p)
    x = _c64(x0).copy()
    xc, yc = _c64(xc), _c64(yc)
    lib.or_sirt_f64(ct.byref(prj_of(prj)), ct.byref(vol_of(vol)), _d(xc), _d(yc), _f(p), _d(x), iters, lam)
    return x


def mace_f64(prjs, vol, xcs, ycs, ps, M, S, lam, rho):
    K = len(prjs)
    xcs = [_c64(a) for a in xcs]
    ycs = [_c64(a) for a in ycs]
    ps = [_c32(a) for a in ps]
    parr = (OrPrj * K)(*[prj_of(p) for p in prjs])
    xa = (ct.POINTER(ct.c_double) * K)(*[_d(a) for a in xcs])
    ya = (ct.POINTER(ct.c_double) * K)(*[_d(a) for a in ycs])
    pa = (ct.POINTER(ct.c_float) * K)(*[_f(a) for a in ps])
    out = np.empty((vol.Z, vol.Y, vol.X), np.float64)
    rc = lib.or_mace_f64(K, parr, ct.byref(vol_of(vol)), xa, ya, pa, M, S, lam, rho, _d(out))
    assert rc == 0
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
