"""Loads libmtr.so (built in-tree) and declares the C ABI prototypes."""
from __future__ import annotations

import ctypes as ct
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmtr.so")


class MtrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class VolumeSize(ct.Structure):
    """Layout of VolumeSize, MrcFileIOBaseFunction.h:57-70."""

    _fields_ = [(n, ct.c_int) for n in ("Xstart", "Xend", "Ystart", "Yend", "Zstart", "Zend", "X", "Y", "Z")]


class ProjectionSize(ct.Structure):
    """Layout of ProjectionSize, MrcFileIOBaseFunction.h:73-78."""

    _fields_ = [("X", ct.c_int), ("Y", ct.c_int), ("AngN", ct.c_int)]


class Slice(ct.Structure):
    """Layout of Slice, MrcFileIOBaseFunction.h:50-54."""

    _fields_ = [("z_start", ct.c_int), ("z_end", ct.c_int)]


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C multi-tilt-reconstuction_amd/csrc). There is no CPU fallback."
    )

lib = ct.CDLL(LIB_PATH)

_i, _d, _f, _v, _i64, _sz = ct.c_int, ct.c_double, ct.c_float, ct.c_void_p, ct.c_int64, ct.c_size_t
_pf, _pd, _pi = ct.POINTER(ct.c_float), ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)
_pvol, _pprj = ct.POINTER(VolumeSize), ct.POINTER(ProjectionSize)

_PROTOS = {
    "mtr_abi_version": (_i, []),
    "mtr_last_error": (ct.c_char_p, []),
    "mtr_ctx_create": (_i, [_i, ct.POINTER(_v)]),
    "mtr_ctx_destroy": (_i, [_v]),
    "mtr_ctx_set_precision": (_i, [_v, _i]),
    "mtr_ctx_use_graphs": (_i, [_v, _i]),
    "mtr_ctx_autotune": (_i, [_v, _i]),
    "mtr_agent_tiles": (_i, [_v, _i, ct.POINTER(_i), ct.POINTER(_i)]),
    "mtr_agent_fp_window": (_i, [_v, _i, ct.POINTER(_i)]),
    "mtr_ctx_set_volume": (_i, [_v, _pvol]),
    "mtr_agent_add": (_i, [_v, _pprj, _pd, _pd, _pf, _pi]),
    "mtr_agent_add_device": (_i, [_v, _pprj, _pd, _pd, _v, _pi]),
    "mtr_agent_count": (_i, [_v, _pi]),
    "mtr_agent_footprint": (_i, [_v, _pprj, ct.POINTER(_sz)]),
    "mtr_agent_device_bytes": (_i, [_v, _i, ct.POINTER(_sz), ct.POINTER(_sz)]),
    "mtr_agent_kernels": (_i, [_v, _i, ct.POINTER(_i), ct.POINTER(_i)]),
    "mtr_bpt": (_i, [_v, _i, _pf]),
    "mtr_sirt": (_i, [_v, _i, _pf, _i, _f, _pf]),
    "mtr_form_z": (_i, [_v, _i]),
    "mtr_mann": (_i, [_v, _i, _f]),
    "mtr_mean": (_i, [_v]),
    "mtr_mace_init": (_i, [_v]),
    "mtr_mace_solve": (_i, [_v, _i, _i, _f, _f, _i, _pf]),
    "mtr_mace_continue": (_i, [_v, _i, _i, _f, _f, _d, _pi, _pf]),
    "mtr_checkpoint_save": (_i, [_v, ct.c_char_p, _i]),
    "mtr_checkpoint_load": (_i, [_v, ct.c_char_p, _pi]),
    "mtr_get_volume": (_i, [_v, _i, _i, _pf]),
    "mtr_set_volume": (_i, [_v, _i, _i, _pf]),
    "mtr_wbar_device": (_i, [_v, ct.POINTER(_v)]),
    "mtr_residual_norm2": (_i, [_v, _i, _pd]),
    "mtr_consensus_diff2": (_i, [_v, _pd]),
    "mtr_op_fp": (_i, [_v, _pprj, _pvol, _pd, _pd, _pf, _pf, _pf]),
    "mtr_op_fp_noalloc": (_i, [_v, _pprj, _pvol, _pd, _pd, _v, _v, _v]),
    "mtr_op_residual": (_i, [_v, _i64, _pf, _pf, _pf, _pf]),
    "mtr_op_bp": (_i, [_v, _pprj, _pvol, _pd, _pd, _pf, _pf, _f]),
    "mtr_op_bpt": (_i, [_v, _pprj, _pvol, _pd, _pd, _pf, _pf]),
    "mtr_divide_slice": (_i, [_i, _i, _i, _i, ct.POINTER(Slice)]),
    "mtr_op_form_z": (_i, [_v, _i64, _pf, _pf, _pf]),
    "mtr_op_mann": (_i, [_v, _i64, _pf, _pf, _pf, _f]),
    "mtr_op_mean": (_i, [_v, _i64, _i, ct.POINTER(_pf), _pf]),
    "mtr_timing_enable": (_i, [_v, _i]),
    "mtr_timing_reset": (_i, [_v]),
    "mtr_timing_get": (_i, [_v, _i, ct.POINTER(_i64), _pd]),
    "mtr_sync": (_i, [_v]),
    "mtr_calibrate_copy": (_i, [_v, _sz, _pd]),
    # request server
    "mtr_server_create": (_i, [_i, _i, _sz, _i, ct.POINTER(_v)]),
    "mtr_server_destroy": (_i, [_v]),
    "mtr_server_footprint": (_i, [_v, _pvol, _i, _pprj, ct.POINTER(_sz)]),
    "mtr_server_submit": (_i, [_v, _pvol, _i, _pprj, ct.POINTER(_pd), ct.POINTER(_pd), ct.POINTER(_pf), _i, _i, _f,
                               _f, ct.POINTER(_i64)]),
    "mtr_server_step": (_i, [_v, _pi]),
    "mtr_server_status": (_i, [_v, _pi, _pi, _pi, ct.POINTER(_sz)]),
    "mtr_server_result": (_i, [_v, _i64, _pf, _pd]),
    "mtr_server_now": (_i, [_v, _pd]),
    "mtr_server_timing": (_i, [_v, _i]),
    "mtr_server_timing_get": (_i, [_v, _i, ct.POINTER(_i64), _pd]),
    # file boundary (mtr_io.h)
    "mtr_mrc_read_header": (_i, [ct.c_char_p, _v]),
    "mtr_mrc_read_data": (_i, [ct.c_char_p, _pf, _i64]),
    "mtr_mrc_write_volume": (_i, [ct.c_char_p, _pvol, _pf]),
    "mtr_mrc_write_stack": (_i, [ct.c_char_p, _i, _i, _i, _pf]),
    "mtr_txbr_read": (_i, [ct.c_char_p, _pvol, _pd, _pd, _i, _pi]),
    "mtr_txbr_write": (_i, [ct.c_char_p, _pvol, _pd, _pd, _i]),
    # synthetic generator (mtr_synth.h)
    "mtr_synth_geometry": (_i, [_pvol, _pprj, _i, _d, _pd, _pd, _pd]),
    "mtr_synth_phantom": (_i, [ct.c_uint64, _pvol, _i, _d, _d, _d, _d, _pd]),
    "mtr_synth_project": (_i, [_pvol, _pprj, _pd, _pd, _i, _pd, _pf, _i]),
    "mtr_synth_noise": (_i, [ct.c_uint64, _i64, _pf, _d, _pd]),
    "mtr_synth_voxelize": (_i, [_pvol, _i, _pd, _pf, _i]),
    "mtr_synth_uniform": (_i, [ct.c_uint64, _i64, _pd]),
    "mtr_synth_normal": (_i, [ct.c_uint64, _i64, _pd]),
}

for _name, (_res, _args) in _PROTOS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_PROTOS)
