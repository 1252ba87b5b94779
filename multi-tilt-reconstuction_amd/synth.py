"""Seeded synthetic problems (phantom, geometry, analytic projections, noise).

Wraps the generator in ``include/mtr_synth.h``.  The named configurations
mirror the workloads the reconstructor is measured on:

  c1  256x256x64, 200 ellipsoids, tilt axis along Z, 61 views -60..+60 deg,
      256x64 detector, 1 agent, 10 SIRT iterations at lambda 1.0
  c2  512x512x128, 400 ellipsoids, in-plane axes 0 / 90 deg, 512x512
      detector, 10 MACE steps of 5 SIRT iterations, lambda 0.5, rho 0.5
  c3  1024x1024x256, 1000 ellipsoids, axes 0/45/90/135 deg, 20 MACE steps
  c4  1024x1024x256, 1000 ellipsoids, axes 0/60/120 deg, 20 MACE steps
  c5  a stream of requests (serving_stream): Poisson arrivals, uneven
      volumes and axis counts
All use Gaussian noise with sigma = 1% of the largest ray sum; agent a of a
problem seeded s draws its noise from seed 100*s + a.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field

import numpy as np

from ._lib import MtrError, ProjectionSize, VolumeSize, lib

GEOM_SLAB, GEOM_AXIS_Z = 0, 1
ELL_DOUBLES = 16


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise MtrError(rc, f"{what}: {lib.mtr_last_error().decode(errors='replace')}")


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _fp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def tilts(lo: float = -60.0, hi: float = 60.0, step: float = 2.0) -> np.ndarray:
    n = int(round((hi - lo) / step)) + 1
    return np.array([lo + step * i for i in range(n)], dtype=np.float64)


def geometry(vol: VolumeSize, prj: ProjectionSize, kind: int, phi_deg: float, theta_deg) -> tuple[np.ndarray, np.ndarray]:
    th = np.ascontiguousarray(theta_deg, dtype=np.float64)
    if th.size != prj.AngN:
        raise ValueError("one tilt per view")
    xc = np.empty(4 * prj.AngN, np.float64)
    yc = np.empty(4 * prj.AngN, np.float64)
    _chk(lib.mtr_synth_geometry(ct.byref(vol), ct.byref(prj), kind, phi_deg, _dp(th), _dp(xc), _dp(yc)), "geometry")
    return xc, yc


def phantom(seed: int, vol: VolumeSize, n: int, amin: float = 4.0, amax: float = 40.0, dmin: float = 0.1,
            dmax: float = 1.0) -> np.ndarray:
    ell = np.empty(max(n, 1) * ELL_DOUBLES, np.float64)
    _chk(lib.mtr_synth_phantom(seed, ct.byref(vol), n, amin, amax, dmin, dmax, _dp(ell)), "phantom")
    return ell[: n * ELL_DOUBLES].reshape(n, ELL_DOUBLES)


def project(vol: VolumeSize, prj: ProjectionSize, xc, yc, ell: np.ndarray, threads: int = 0) -> np.ndarray:
    ell = np.ascontiguousarray(ell, dtype=np.float64)
    p = np.empty((prj.AngN, prj.Y, prj.X), np.float32)
    n = ell.shape[0]
    e = ell if n else np.zeros(ELL_DOUBLES)
    _chk(lib.mtr_synth_project(ct.byref(vol), ct.byref(prj), _dp(np.ascontiguousarray(xc, np.float64)),
                               _dp(np.ascontiguousarray(yc, np.float64)), n, _dp(e), _fp(p), threads), "project")
    return p


def add_noise(seed: int, p: np.ndarray, frac: float = 0.01) -> float:
    assert p.dtype == np.float32 and p.flags.c_contiguous
    sig = ct.c_double()
    _chk(lib.mtr_synth_noise(seed, p.size, _fp(p), frac, ct.byref(sig)), "noise")
    return sig.value


def voxelize(vol: VolumeSize, ell: np.ndarray, threads: int = 0) -> np.ndarray:
    ell = np.ascontiguousarray(ell, dtype=np.float64)
    x = np.empty((vol.Z, vol.Y, vol.X), np.float32)
    n = ell.shape[0]
    e = ell if n else np.zeros(ELL_DOUBLES)
    _chk(lib.mtr_synth_voxelize(ct.byref(vol), n, _dp(e), _fp(x), threads), "voxelize")
    return x


def uniform(seed: int, n: int) -> np.ndarray:
    o = np.empty(n, np.float64)
    _chk(lib.mtr_synth_uniform(seed, n, _dp(o)), "uniform")
    return o


def normal(seed: int, n: int) -> np.ndarray:
    o = np.empty(n, np.float64)
    _chk(lib.mtr_synth_normal(seed, n, _dp(o)), "normal")
    return o


@dataclass
class Problem:
    """A reconstruction request: shared box, K tilt series, solve parameters."""

    name: str
    vol: VolumeSize
    prjs: list = field(default_factory=list)
    xcs: list = field(default_factory=list)
    ycs: list = field(default_factory=list)
    ps: list = field(default_factory=list)
    M: int = 1
    S: int = 10
    lam: float = 1.0
    rho: float = 0.5
    seed: int = 0
    n_ellipsoids: int = 0

    @property
    def K(self) -> int:
        return len(self.prjs)

    @property
    def N(self) -> int:
        return self.vol.X * self.vol.Y * self.vol.Z

    @property
    def P(self) -> list:
        return [p.X * p.Y * p.AngN for p in self.prjs]


def make_problem(name: str, X: int, Y: int, Z: int, n_ell: int, phis, *, kind: int = GEOM_SLAB, nu: int | None = None,
                 nv: int | None = None, theta=None, seed: int = 1, noise: float = 0.01, M: int = 1, S: int = 10,
                 lam: float = 1.0, rho: float = 0.5, amin: float = 4.0, amax: float = 40.0,
                 threads: int = 0) -> Problem:
    vol = VolumeSize(0, X, 0, Y, 0, Z, X, Y, Z)
    th = tilts() if theta is None else np.asarray(theta, np.float64)
    if kind == GEOM_AXIS_Z:
        nu = X if nu is None else nu
        nv = Z if nv is None else nv
    else:
        nu = X if nu is None else nu
        nv = Y if nv is None else nv
    prj = ProjectionSize(nu, nv, th.size)
    ell = phantom(seed, vol, n_ell, amin, amax)
    pb = Problem(name, vol, M=M, S=S, lam=lam, rho=rho, seed=seed, n_ellipsoids=n_ell)
    for a, phi in enumerate(phis):
        xc, yc = geometry(vol, prj, kind, float(phi), th)
        p = project(vol, prj, xc, yc, ell, threads)
        if noise > 0:
            add_noise(seed * 100 + a, p, noise)
        pb.prjs.append(ProjectionSize(prj.X, prj.Y, prj.AngN))
        pb.xcs.append(xc)
        pb.ycs.append(yc)
        pb.ps.append(p)
    return pb


CONFIGS = {
    "c1": dict(X=256, Y=256, Z=64, n_ell=200, phis=[0.0], kind=GEOM_AXIS_Z, nu=256, nv=64, seed=1, M=1, S=10,
               lam=1.0, rho=0.5),
    "c2": dict(X=512, Y=512, Z=128, n_ell=400, phis=[0.0, 90.0], seed=2, M=10, S=5, lam=0.5, rho=0.5),
    "c3": dict(X=1024, Y=1024, Z=256, n_ell=1000, phis=[0.0, 45.0, 90.0, 135.0], seed=3, M=20, S=10, lam=0.5,
               rho=0.5),
    "c4": dict(X=1024, Y=1024, Z=256, n_ell=1000, phis=[0.0, 60.0, 120.0], seed=4, M=20, S=10, lam=0.5, rho=0.5),
}


# Serving stream: 64 requests with Poisson arrivals at mean 4 per second;
# each draws its volume xy, z and axis count uniformly, axes evenly spaced
# over 180 degrees, 61 views over +-60 degrees, 10 MACE steps of 5 SIRT.
SERVING = dict(n=64, rate=4.0, seed=5, xy=(256, 384, 512, 768), z=(64, 128, 256), max_axes=4, M=10, S=5,
               lam=0.5, rho=0.5, n_ell=24)


def serving_params(q: int, n: int = SERVING["n"], rate: float = SERVING["rate"], seed: int = SERVING["seed"],
                   xy=SERVING["xy"], z=SERVING["z"], max_axes: int = SERVING["max_axes"]):
    """(arrival seconds, xy, z, axis count) of request q of the seeded stream:
    exponential inter-arrival gaps (-ln(1 - U) / rate) and uniform draws of
    xy, z and the axis count from the generator's U(0,1) sequence of `seed`."""
    u = uniform(seed, 4 * n)
    t = 0.0
    for k in range(q + 1):
        t += -np.log1p(-u[4 * k]) / rate
    s_xy = xy[min(int(u[4 * q + 1] * len(xy)), len(xy) - 1)]
    s_z = z[min(int(u[4 * q + 2] * len(z)), len(z) - 1)]
    K = 1 + min(int(u[4 * q + 3] * max_axes), max_axes - 1)
    return t, s_xy, s_z, K


def serving_request(q: int, n: int = SERVING["n"], rate: float = SERVING["rate"], seed: int = SERVING["seed"],
                    xy=SERVING["xy"], z=SERVING["z"], max_axes: int = SERVING["max_axes"], M: int = SERVING["M"],
                    S: int = SERVING["S"], n_ell: int = SERVING["n_ell"], theta=None, threads: int = 0):
    """(arrival seconds, Problem) of request q of the seeded stream: K axes
    evenly spaced over 180 degrees, seed 1000 seed + q."""
    t, s_xy, s_z, K = serving_params(q, n, rate, seed, xy, z, max_axes)
    phis = [180.0 * a / K for a in range(K)]
    pb = make_problem(f"req{q}", s_xy, s_xy, s_z, n_ell, phis, seed=seed * 1000 + q, M=M, S=S,
                      lam=SERVING["lam"], rho=SERVING["rho"], theta=theta, threads=threads)
    return t, pb


def serving_stream(n: int = SERVING["n"], rate: float = SERVING["rate"], seed: int = SERVING["seed"],
                   xy=SERVING["xy"], z=SERVING["z"], max_axes: int = SERVING["max_axes"], M: int = SERVING["M"],
                   S: int = SERVING["S"], n_ell: int = SERVING["n_ell"], theta=None, threads: int = 0):
    """[(arrival seconds, Problem)] of the seeded request stream (serving_request for q = 0 .. n - 1)."""
    return [serving_request(q, n, rate, seed, xy, z, max_axes, M, S, n_ell, theta, threads) for q in range(n)]


def config(name: str, threads: int = 0, **overrides) -> Problem:
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    return make_problem(name, threads=threads, **kw)
