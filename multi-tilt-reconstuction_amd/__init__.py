"""MI355X-native multi-axis tomography reconstructor: Python binding.

A thin ctypes layer over the C ABI in ``include/mtr.h`` (``libmtr.so``, built
in-tree by ``csrc/Makefile``).  Nothing here computes: every operation runs
in the HIP library, and importing the package fails loudly when the library
has not been built.

Reference correspondence (Autumn1998/Multi-tilt-Reconstuction):
  Context.bpt            ComputeInitialModel   FunctionOnGPU.cu:138-179
  Context.sirt           runSIRTOnGPU          FunctionOnGPU.cu:182-266
  Context.form_z         initialZdata          FunctionOnGPU.cu:291-313
  Context.mann           runMaceOnCPU          MultiAxis.cpp:33-38
  Context.mean           consensus average     MultiAxis.cpp:61-66
  Context.mace_init      BPT-sum initial model MultiAxis.cpp:163-200
  Context.mace_solve     reconstructModel      MultiAxis.cpp:40-71
  Context.op_fp          computeDivisor/Reproj FunctionOnGPU.cu:316-333, 27-52
  Context.op_residual    residual loop         FunctionOnGPU.cu:228-236
  Context.op_bp          sirtBackProjOnGPU     FunctionOnGPU.cu:111-135
  Context.op_bpt         backProjOnGPU         FunctionOnGPU.cu:82-109
  divide_slice           divideSlice           MultiAxis.cpp:13-31
"""
from __future__ import annotations

import ctypes as _ct
import os as _os

import numpy as _np

from . import mrc, synth  # noqa: F401  (file boundary and generator helpers)
from ._lib import (  # noqa: F401
    LIB_PATH,
    MtrError,
    ProjectionSize,
    Slice,
    VolumeSize,
    lib,
)

PREC_FP32, PREC_BF16, PREC_FP16 = 0, 1, 2
ADMIT_ARRIVAL, ADMIT_SMALLEST = 0, 1
FPK_GENERAL, FPK_STREAM, FPK_COLUMN = 0, 1, 2
BPK_GENERAL, BPK_SEPARABLE, BPK_COLUMN = 0, 1, 2
FORM_FAST, FORM_PLAIN = 0, 1
VOL_WORK, VOL_W, VOL_WBAR = 0, 1, 2
K_FP, K_BP, K_BPT, K_CONSENSUS, K_NORMALISER, K_ELEMENTWISE, K_SPLAT = range(7)
# per-kernel parts of the FP and BP classes (SIRT launches only)
K_FP_STREAM, K_FP_COLUMN, K_FP_COLUMN_RES, K_FP_GENERAL, K_BP_SEPARABLE, K_BP_COLUMN, K_BP_GENERAL = range(7, 14)
K_FP_STREAM_RES = 14  # Z-segmented streaming FP: segment sums + residual
KERNEL_NAMES = ["fp", "bp", "bpt", "consensus", "normaliser", "elementwise", "splat", "fp_stream", "fp_column",
                "fp_column_res", "fp_general", "bp_separable", "bp_column", "bp_general", "fp_stream_res"]


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib.mtr_last_error().decode(errors="replace")
        raise MtrError(rc, f"{what}: {msg}")


def _f32(a, shape=None) -> _np.ndarray:
    a = _np.ascontiguousarray(a, dtype=_np.float32)
    if shape is not None and a.size != int(_np.prod(shape)):
        raise ValueError(f"expected {int(_np.prod(shape))} elements, got {a.size}")
    return a


def _f64(a) -> _np.ndarray:
    return _np.ascontiguousarray(a, dtype=_np.float64)


def _fp(a: _np.ndarray):
    return a.ctypes.data_as(_ct.POINTER(_ct.c_float))


def _dp(a: _np.ndarray):
    return a.ctypes.data_as(_ct.POINTER(_ct.c_double))


def volume_size(X: int, Y: int, Z: int, Xstart: int = 0, Ystart: int = 0, Zstart: int = 0) -> VolumeSize:
    """Box [start, start+size) per axis (VolumeSize, MrcFileIOBaseFunction.h:57-70)."""
    return VolumeSize(Xstart, Xstart + X, Ystart, Ystart + Y, Zstart, Zstart + Z, X, Y, Z)


def projection_size(nu: int, nv: int, A: int) -> ProjectionSize:
    return ProjectionSize(nu, nv, A)


def divide_slice(myid: int, nparts: int, Z: int, Zstart: int = 0) -> tuple[int, int]:
    s = Slice()
    _check(lib.mtr_divide_slice(myid, nparts, Z, Zstart, _ct.byref(s)), "divide_slice")
    return s.z_start, s.z_end


class Context:
    """One device context: resident agents, their MACE state and w-bar."""

    def __init__(self, device: int = 0, precision: int = PREC_FP32):
        self._h = _ct.c_void_p()
        _check(lib.mtr_ctx_create(device, _ct.byref(self._h)), "mtr_ctx_create")
        if precision != PREC_FP32:
            _check(lib.mtr_ctx_set_precision(self._h, precision), "set_precision")
        self.vol: VolumeSize | None = None
        self.prjs: list[ProjectionSize] = []

    # -- lifetime --------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib.mtr_ctx_destroy(self._h)
            self._h = _ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -------------------------------------------------------------
    def autotune(self, on: bool) -> None:
        _check(lib.mtr_ctx_autotune(self._h, int(on)), "autotune")

    def agent_tiles(self, agent: int) -> tuple[int, int]:
        """(run length, lane axis) of an agent's back-projection tiles."""
        r, l = _ct.c_int(), _ct.c_int()
        _check(lib.mtr_agent_tiles(self._h, agent, _ct.byref(r), _ct.byref(l)), "agent_tiles")
        return r.value, l.value

    def agent_fp_window(self, agent: int) -> int:
        """LDS row slots per wave of the column-plan FP (0: whole rows)."""
        w = _ct.c_int()
        _check(lib.mtr_agent_fp_window(self._h, agent, _ct.byref(w)), "agent_fp_window")
        return w.value

    def use_graphs(self, on: bool) -> None:
        _check(lib.mtr_ctx_use_graphs(self._h, int(on)), "use_graphs")

    def set_volume(self, vol: VolumeSize) -> None:
        _check(lib.mtr_ctx_set_volume(self._h, _ct.byref(vol)), "set_volume")
        self.vol = vol

    @property
    def shape(self) -> tuple[int, int, int]:
        assert self.vol is not None
        return (self.vol.Z, self.vol.Y, self.vol.X)

    def add_agent(self, prj: ProjectionSize, xc, yc, p) -> int:
        xc, yc = _f64(xc), _f64(yc)
        if xc.size != 4 * prj.AngN or yc.size != 4 * prj.AngN:
            raise ValueError("xc/yc must hold 4*AngN coefficients")
        p = _f32(p, (prj.AngN, prj.Y, prj.X))
        aid = _ct.c_int()
        _check(lib.mtr_agent_add(self._h, _ct.byref(prj), _dp(xc), _dp(yc), _fp(p), _ct.byref(aid)), "agent_add")
        self.prjs.append(prj)
        return aid.value

    def add_agent_device(self, prj: ProjectionSize, xc, yc, p_dev_ptr: int) -> int:
        xc, yc = _f64(xc), _f64(yc)
        aid = _ct.c_int()
        _check(
            lib.mtr_agent_add_device(self._h, _ct.byref(prj), _dp(xc), _dp(yc), _ct.c_void_p(p_dev_ptr), _ct.byref(aid)),
            "agent_add_device",
        )
        self.prjs.append(prj)
        return aid.value

    def agent_footprint(self, prj: ProjectionSize) -> int:
        b = _ct.c_size_t()
        _check(lib.mtr_agent_footprint(self._h, _ct.byref(prj), _ct.byref(b)), "agent_footprint")
        return b.value

    def agent_device_bytes(self, agent: int) -> tuple[int, int]:
        """(bytes held now, most bytes held while the agent was added)."""
        r, pk = _ct.c_size_t(), _ct.c_size_t()
        _check(lib.mtr_agent_device_bytes(self._h, agent, _ct.byref(r), _ct.byref(pk)), "agent_device_bytes")
        return r.value, pk.value

    def agent_kernels(self, agent: int) -> tuple[int, int]:
        """(forward, back) projection kernels chosen for an agent: FPK_* and BPK_*."""
        f, b = _ct.c_int(), _ct.c_int()
        _check(lib.mtr_agent_kernels(self._h, agent, _ct.byref(f), _ct.byref(b)), "agent_kernels")
        return f.value, b.value

    # -- hot path ----------------------------------------------------------
    def bpt(self, agent: int, out: bool = True):
        o = _np.empty(self.shape, _np.float32) if out else None
        _check(lib.mtr_bpt(self._h, agent, _fp(o) if out else None), "bpt")
        return o

    def sirt(self, agent: int, iters: int, lam: float, x0=None, out: bool = True):
        x0p = None
        if x0 is not None:
            x0 = _f32(x0, self.shape)
            x0p = _fp(x0)
        o = _np.empty(self.shape, _np.float32) if out else None
        _check(lib.mtr_sirt(self._h, agent, x0p, iters, _ct.c_float(lam), _fp(o) if out else None), "sirt")
        return o

    def form_z(self, agent: int) -> None:
        _check(lib.mtr_form_z(self._h, agent), "form_z")

    def mann(self, agent: int, rho: float) -> None:
        _check(lib.mtr_mann(self._h, agent, _ct.c_float(rho)), "mann")

    def mean(self) -> None:
        _check(lib.mtr_mean(self._h), "mean")

    def mace_init(self) -> None:
        _check(lib.mtr_mace_init(self._h), "mace_init")

    def mace_solve(self, M: int, S: int, lam: float, rho: float, form: int = FORM_FAST, out: bool = True):
        o = _np.empty(self.shape, _np.float32) if out else None
        _check(
            lib.mtr_mace_solve(self._h, M, S, _ct.c_float(lam), _ct.c_float(rho), form, _fp(o) if out else None),
            "mace_solve",
        )
        return o

    def mace_continue(self, M: int, S: int, lam: float, rho: float, rel_tol: float = 0.0, out: bool = True):
        """At most M more fast MACE steps from the current state (after mace_init, a solve, or a checkpoint
        restored with set_volume_data); stops early once the FP64 consensus difference changes by at most
        rel_tol of its previous value.  Returns (w-bar or None, steps run)."""
        o = _np.empty(self.shape, _np.float32) if out else None
        n = _ct.c_int()
        _check(lib.mtr_mace_continue(self._h, M, S, _ct.c_float(lam), _ct.c_float(rho), _ct.c_double(rel_tol),
                                     _ct.byref(n), _fp(o) if out else None), "mace_continue")
        return o, n.value

    def checkpoint_save(self, path: str, steps_done: int) -> None:
        """Writes the MACE state (every w_a, then w-bar) with the caller's count of steps done."""
        _check(lib.mtr_checkpoint_save(self._h, _os.fsencode(path), steps_done), "checkpoint_save")

    def checkpoint_load(self, path: str) -> int:
        """Replaces the MACE state by a checkpoint of the same inputs; returns its count of steps done."""
        n = _ct.c_int()
        _check(lib.mtr_checkpoint_load(self._h, _os.fsencode(path), _ct.byref(n)), "checkpoint_load")
        return n.value

    def get_volume(self, which: int, agent: int = 0) -> _np.ndarray:
        o = _np.empty(self.shape, _np.float32)
        _check(lib.mtr_get_volume(self._h, which, agent, _fp(o)), "get_volume")
        return o

    def set_volume_data(self, which: int, agent: int, data) -> None:
        d = _f32(data, self.shape)
        _check(lib.mtr_set_volume(self._h, which, agent, _fp(d)), "set_volume_data")

    def residual_norm2(self, agent: int) -> float:
        v = _ct.c_double()
        _check(lib.mtr_residual_norm2(self._h, agent, _ct.byref(v)), "residual_norm2")
        return v.value

    def consensus_diff2(self) -> float:
        v = _ct.c_double()
        _check(lib.mtr_consensus_diff2(self._h, _ct.byref(v)), "consensus_diff2")
        return v.value

    def calibrate_copy(self, nbytes: int) -> float:
        ms = _ct.c_double()
        _check(lib.mtr_calibrate_copy(self._h, nbytes, _ct.byref(ms)), "calibrate_copy")
        return ms.value

    def sync(self) -> None:
        _check(lib.mtr_sync(self._h), "sync")

    # -- timing ------------------------------------------------------------
    def timing(self, on: bool) -> None:
        _check(lib.mtr_timing_enable(self._h, 1 if on else 0), "timing_enable")

    def timing_reset(self) -> None:
        _check(lib.mtr_timing_reset(self._h), "timing_reset")

    def timing_get(self, kernel: int) -> tuple[int, float]:
        n, ms = _ct.c_int64(), _ct.c_double()
        _check(lib.mtr_timing_get(self._h, kernel, _ct.byref(n), _ct.byref(ms)), "timing_get")
        return n.value, ms.value

    # -- stateless operators -----------------------------------------------
    def op_fp(self, prj: ProjectionSize, vol: VolumeSize, xc, yc, x):
        xc, yc = _f64(xc), _f64(yc)
        x = _f32(x, (vol.Z, vol.Y, vol.X))
        s = _np.empty((prj.AngN, prj.Y, prj.X), _np.float32)
        c = _np.empty_like(s)
        _check(
            lib.mtr_op_fp(self._h, _ct.byref(prj), _ct.byref(vol), _dp(xc), _dp(yc), _fp(x), _fp(s), _fp(c)), "op_fp"
        )
        return s, c

    def op_fp_noalloc(self, prj: ProjectionSize, vol: VolumeSize, xc, yc, x_dev: int, s_dev: int, c_dev: int) -> None:
        """Forward projection on device buffers (addresses as ints) without
        any device allocation; see mtr_op_fp_noalloc in include/mtr.h."""
        xc, yc = _f64(xc), _f64(yc)
        _check(lib.mtr_op_fp_noalloc(self._h, _ct.byref(prj), _ct.byref(vol), _dp(xc), _dp(yc), x_dev, s_dev, c_dev),
               "op_fp_noalloc")

    def op_residual(self, p, s, c):
        p, s, c = _f32(p), _f32(s), _f32(c)
        r = _np.empty_like(p)
        _check(lib.mtr_op_residual(self._h, p.size, _fp(p), _fp(s), _fp(c), _fp(r)), "op_residual")
        return r

    def op_form_z(self, w, wbar):
        w, wbar = _f32(w), _f32(wbar)
        z = _np.empty_like(w)
        _check(lib.mtr_op_form_z(self._h, w.size, _fp(w), _fp(wbar), _fp(z)), "op_form_z")
        return z

    def op_mann(self, v, w, z, rho: float):
        v, z = _f32(v), _f32(z)
        wo = _f32(w).copy()
        _check(lib.mtr_op_mann(self._h, wo.size, _fp(v), _fp(wo), _fp(z), _ct.c_float(rho)), "op_mann")
        return wo

    def op_mean(self, ws):
        ws = [_f32(w) for w in ws]
        arr = (_ct.POINTER(_ct.c_float) * len(ws))(*[_fp(w) for w in ws])
        out = _np.empty_like(ws[0])
        _check(lib.mtr_op_mean(self._h, ws[0].size, len(ws), arr, _fp(out)), "op_mean")
        return out

    def op_bp(self, prj: ProjectionSize, vol: VolumeSize, xc, yc, r, x, lam: float):
        xc, yc = _f64(xc), _f64(yc)
        r = _f32(r, (prj.AngN, prj.Y, prj.X))
        xo = _f32(x, (vol.Z, vol.Y, vol.X)).copy()
        _check(
            lib.mtr_op_bp(self._h, _ct.byref(prj), _ct.byref(vol), _dp(xc), _dp(yc), _fp(r), _fp(xo), _ct.c_float(lam)),
            "op_bp",
        )
        return xo

    def op_bpt(self, prj: ProjectionSize, vol: VolumeSize, xc, yc, p):
        xc, yc = _f64(xc), _f64(yc)
        p = _f32(p, (prj.AngN, prj.Y, prj.X))
        xo = _np.empty((vol.Z, vol.Y, vol.X), _np.float32)
        _check(lib.mtr_op_bpt(self._h, _ct.byref(prj), _ct.byref(vol), _dp(xc), _dp(yc), _fp(p), _fp(xo)), "op_bpt")
        return xo


class Server:
    """Continuous-batching request server on one device (mtr_server_*).

    submit() copies a request (a synth.Problem, or its parts) into the
    server; step() admits pending requests that fit the HBM budget, runs one
    MACE step of every active request with their SIRT iterations batched
    into shared launches, and retires the finished ones."""

    def __init__(self, device: int = 0, precision: int = PREC_FP32, hbm_budget: int = 200 << 30,
                 max_active: int = 64):
        self._h = _ct.c_void_p()
        _check(lib.mtr_server_create(device, precision, hbm_budget, max_active, _ct.byref(self._h)),
               "mtr_server_create")
        self.shapes: dict[int, tuple[int, int, int]] = {}

    def close(self) -> None:
        if self._h:
            lib.mtr_server_destroy(self._h)
            self._h = _ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arrays(prjs):
        return (ProjectionSize * len(prjs))(*[ProjectionSize(p.X, p.Y, p.AngN) for p in prjs])

    def footprint(self, vol: VolumeSize, prjs) -> int:
        b = _ct.c_size_t()
        _check(lib.mtr_server_footprint(self._h, _ct.byref(vol), len(prjs), self._arrays(prjs), _ct.byref(b)),
               "server_footprint")
        return b.value

    def submit(self, vol: VolumeSize, prjs, xcs, ycs, ps, M: int, S: int, lam: float, rho: float) -> int:
        K = len(prjs)
        xcs = [_f64(x) for x in xcs]
        ycs = [_f64(y) for y in ycs]
        ps = [_f32(p, (pr.AngN, pr.Y, pr.X)) for p, pr in zip(ps, prjs)]
        PD = _ct.POINTER(_ct.c_double)
        PF = _ct.POINTER(_ct.c_float)
        xa = (PD * K)(*[_dp(x) for x in xcs])
        ya = (PD * K)(*[_dp(y) for y in ycs])
        pa = (PF * K)(*[_fp(p) for p in ps])
        rid = _ct.c_int64()
        _check(lib.mtr_server_submit(self._h, _ct.byref(vol), K, self._arrays(prjs), xa, ya, pa, M, S, lam, rho,
                                     _ct.byref(rid)), "server_submit")
        self.shapes[rid.value] = (vol.Z, vol.Y, vol.X)
        return rid.value

    def submit_problem(self, pb) -> int:
        return self.submit(pb.vol, pb.prjs, pb.xcs, pb.ycs, pb.ps, pb.M, pb.S, pb.lam, pb.rho)

    def set_admission(self, policy: int, max_wait: float = 0.0) -> None:
        """ADMIT_ARRIVAL (oldest first) or ADMIT_SMALLEST (least work first,
        except that a request waiting longer than max_wait seconds goes
        first, in arrival order)."""
        _check(lib.mtr_server_set_admission(self._h, int(policy), _ct.c_double(max_wait)), "server_set_admission")

    def step(self) -> int:
        n = _ct.c_int()
        _check(lib.mtr_server_step(self._h, _ct.byref(n)), "server_step")
        return n.value

    def status(self) -> dict:
        p, a, d = _ct.c_int(), _ct.c_int(), _ct.c_int()
        u = _ct.c_size_t()
        _check(lib.mtr_server_status(self._h, _ct.byref(p), _ct.byref(a), _ct.byref(d), _ct.byref(u)),
               "server_status")
        return {"pending": p.value, "active": a.value, "done": d.value, "hbm_in_use": u.value}

    def result(self, rid: int):
        """(w-bar, (submitted, admitted, completed) seconds)."""
        out = _np.empty(self.shapes[rid], _np.float32)
        t = (_ct.c_double * 3)()
        _check(lib.mtr_server_result(self._h, rid, _fp(out), t), "server_result")
        return out, tuple(t)

    def times(self, rid: int):
        t = (_ct.c_double * 3)()
        lib.mtr_server_result(self._h, rid, None, t)
        return tuple(t)

    def now(self) -> float:
        t = _ct.c_double()
        _check(lib.mtr_server_now(self._h, _ct.byref(t)), "server_now")
        return t.value

    def timing(self, on: bool) -> None:
        _check(lib.mtr_server_timing(self._h, int(on)), "server_timing")

    def timing_get(self, kernel: int) -> tuple[int, float]:
        n = _ct.c_int64()
        ms = _ct.c_double()
        _check(lib.mtr_server_timing_get(self._h, kernel, _ct.byref(n), _ct.byref(ms)), "server_timing_get")
        return n.value, ms.value


def gpu_available() -> bool:
    """True when the library can open a context on device 0."""
    h = _ct.c_void_p()
    if lib.mtr_ctx_create(0, _ct.byref(h)) != 0:
        return False
    lib.mtr_ctx_destroy(h)
    return True


__all__ = [
    "Context",
    "MtrError",
    "VolumeSize",
    "ProjectionSize",
    "Slice",
    "volume_size",
    "projection_size",
    "divide_slice",
    "gpu_available",
    "synth",
    "LIB_PATH",
]
