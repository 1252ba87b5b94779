"""CPU-side checks of the C ABI library and the synthetic generator.

No compute kernel runs here: the library must load, export every symbol the
public headers declare, reject bad arguments with negative status codes, and
its host-only parts (slab partition, generator) must be deterministic and
agree with closed forms.
"""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mtr_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header", ["mtr.h", "mtr_synth.h", "mtr_io.h"])
def test_every_declared_symbol_is_exported(mtr, header):
    names = declared(header)
    assert names, header
    lib = ct.CDLL(mtr.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_error_string(mtr):
    assert mtr.lib.mtr_abi_version() == 1
    s = mtr.Slice()
    assert mtr.lib.mtr_divide_slice(0, 0, 10, 0, ct.byref(s)) == -1
    assert b"partition" in mtr.lib.mtr_last_error()


def test_no_device_is_an_error_not_a_crash(mtr):
    if mtr.gpu_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(mtr.MtrError) as ei:
        mtr.Context(0)
    assert ei.value.code == -8


def test_struct_layouts(mtr):
    assert ct.sizeof(mtr.VolumeSize) == 9 * 4
    assert ct.sizeof(mtr.ProjectionSize) == 3 * 4
    assert ct.sizeof(mtr.Slice) == 2 * 4


def _header_fields(struct):
    txt = open(os.path.join(ROOT, "include", "mtr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + struct + r"\s*;", txt).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            names += [n.strip() for n in decl.split(None, 1)[1].split(",")]
    return names


def _stub_fields(cls):
    """Field names of a ctypes class in the INTEGRATION.md stub."""
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    stub = re.search(r"```python\n(.*?)```", txt, flags=re.S).group(1)
    m = re.search(r"class " + cls + r"\(ct\.Structure\):.*?\n\s+_fields_ = (.*?)\n", stub, flags=re.S)
    assert m, cls
    return re.findall(r'"([A-Za-z_]+)"', m.group(1))


def test_integration_stub_matches_the_abi_layout(mtr):
    # the documented binding must pass the same box as the library's own
    # (VolumeSize field order, MrcFileIOBaseFunction.h:57-70)
    for cls, struct, ours in (("VolumeSize", "mtr_volume_size", mtr.VolumeSize),
                              ("ProjectionSize", "mtr_projection_size", mtr.ProjectionSize)):
        stub = _stub_fields(cls)
        assert stub == [f[0] for f in ours._fields_], cls
        assert stub == _header_fields(struct), cls


def test_divide_slice_matches_reference_rule(mtr):
    # first Z % N parts carry one extra slice (MultiAxis.cpp:13-31)
    assert [mtr.divide_slice(i, 4, 10, 5) for i in range(4)] == [(5, 8), (8, 11), (11, 13), (13, 15)]
    assert [mtr.divide_slice(i, 3, 2, 0) for i in range(3)] == [(0, 1), (1, 2), (2, 2)]


def test_rng_is_deterministic_and_uniform(mtr):
    a = mtr.synth.uniform(42, 100000)
    b = mtr.synth.uniform(42, 100000)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, mtr.synth.uniform(43, 100000))
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005 and abs(a.var() - 1 / 12) < 0.002
    g = mtr.synth.normal(5, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


def test_geometry_maps_centre_to_detector_centre(mtr):
    vol = mtr.volume_size(40, 30, 20, 3, -5, 7)
    prj = mtr.projection_size(50, 36, 5)
    th = np.array([-60.0, -30.0, 0.0, 30.0, 60.0])
    for kind, phi in ((mtr.synth.GEOM_SLAB, 0.0), (mtr.synth.GEOM_SLAB, 37.0), (mtr.synth.GEOM_AXIS_Z, 0.0)):
        xc, yc = mtr.synth.geometry(vol, prj, kind, phi, th)
        C = np.array([(3 + 42) / 2, (-5 + 24) / 2, (7 + 26) / 2])
        for a in range(5):
            u = xc[4 * a] + xc[4 * a + 1: 4 * a + 4] @ C - 1
            v = yc[4 * a] + yc[4 * a + 1: 4 * a + 4] @ C - 1
            assert abs(u - 24.5) < 1e-9 and abs(v - 17.5) < 1e-9
            gu, gv = xc[4 * a + 1: 4 * a + 4], yc[4 * a + 1: 4 * a + 4]
            assert abs(np.linalg.norm(gu) - 1) < 1e-12 and abs(np.linalg.norm(gv) - 1) < 1e-12
            assert abs(gu @ gv) < 1e-12


def test_axis_aligned_geometry_is_exact(mtr):
    vol = mtr.volume_size(32, 32, 16)
    prj = mtr.projection_size(32, 32, 3)
    xc, yc = mtr.synth.geometry(vol, prj, mtr.synth.GEOM_SLAB, 90.0, [-60.0, 0.0, 60.0])
    # phi = 90: v = cv - dX exactly, u independent of X exactly
    assert yc[1::4].tolist() == [-1.0] * 3 and yc[2::4].tolist() == [0.0] * 3
    assert xc[1::4].tolist() == [0.0] * 3
    assert all(float(q).is_integer() for q in yc[0::4])


def test_analytic_projection_of_a_sphere(mtr):
    vol = mtr.volume_size(40, 40, 40)
    prj = mtr.projection_size(40, 40, 2)
    xc, yc = mtr.synth.geometry(vol, prj, mtr.synth.GEOM_SLAB, 0.0, [0.0, 30.0])
    R = 10.0
    ell = np.zeros((1, 16))
    ell[0, :3] = [19.5, 19.5, 19.5]
    ell[0, 3:6] = R
    ell[0, 6:15] = np.eye(3).ravel()
    ell[0, 15] = 2.0
    p = mtr.synth.project(vol, prj, xc, yc, ell)
    i, k = np.meshgrid(np.arange(40), np.arange(40))
    d2 = (i - 19.5) ** 2 + (k - 19.5) ** 2
    expect = np.where(d2 < R * R, 2.0 * 2 * np.sqrt(np.maximum(R * R - d2, 0)), 0.0)
    for a in range(2):
        np.testing.assert_allclose(p[a], expect, atol=1e-4)


def test_voxelized_ellipsoid_volume(mtr):
    vol = mtr.volume_size(64, 64, 64)
    ell = mtr.synth.phantom(3, vol, 1, 10.0, 20.0, 1.0, 1.0)
    ell[0, :3] = 31.5
    x = mtr.synth.voxelize(vol, ell)
    a, b, c = ell[0, 3:6]
    assert abs(x.sum() / (4 / 3 * np.pi * a * b * c) - 1) < 0.03


def test_noise_sigma_and_determinism(mtr):
    p = np.full(100000, 10.0, np.float32)
    q = p.copy()
    s1 = mtr.synth.add_noise(7, p, 0.01)
    s2 = mtr.synth.add_noise(7, q, 0.01)
    assert s1 == pytest.approx(0.1) and np.array_equal(p, q)
    assert abs((p - 10).std() - 0.1) < 0.002


def test_config_shapes(mtr):
    c1 = mtr.synth.CONFIGS["c1"]
    assert (c1["X"], c1["Y"], c1["Z"], c1["nu"], c1["nv"]) == (256, 256, 64, 256, 64)
    assert mtr.synth.tilts().size == 61
    assert mtr.synth.CONFIGS["c2"]["phis"] == [0.0, 90.0]


def test_compat_symbols_are_exported(mtr):
    # the reference-shaped C++ host interface (include/mtr_compat.h)
    out = subprocess.run(["nm", "-DC", mtr.LIB_PATH], capture_output=True, text=True, check=True).stdout
    for name in ("ComputeInitialModel", "runSIRTOnGPU", "initialZdata", "runMaceOnGPU", "runMaceOnCPU",
                 "reconstructModel", "divideSlice", "synchronizeVol", "readOptions", "readMrcHeader",
                 "readMrcAndTXBR", "updataAndWriteModel"):
        assert f"mtr_compat::{name}(" in out, name


def test_serving_stream_is_seeded_and_shaped(mtr):
    a = mtr.synth.serving_stream(n=12, xy=(16, 24), z=(8, 12), theta=[-30.0, 0.0, 30.0])
    b = mtr.synth.serving_stream(n=12, xy=(16, 24), z=(8, 12), theta=[-30.0, 0.0, 30.0])
    assert [t for t, _ in a] == [t for t, _ in b]
    assert all(np.array_equal(p.ps[0], q.ps[0]) for (_, p), (_, q) in zip(a, b))
    times = [t for t, _ in a]
    assert all(t1 > t0 for t0, t1 in zip(times, times[1:]))  # arrivals increase
    for _, pb in a:
        assert pb.vol.X == pb.vol.Y and pb.vol.X in (16, 24) and pb.vol.Z in (8, 12)
        assert 1 <= pb.K <= 4 and pb.M == 10 and pb.S == 5
        # axes evenly spaced over 180 degrees: the v gradient turns by 180/K
        ang = [np.degrees(np.arctan2(-yc[1], yc[2])) % 180.0 for yc in pb.ycs]
        assert np.allclose(np.diff(ang), 180.0 / pb.K) if pb.K > 1 else True
    # one request of the stream on its own is the stream's request
    for q in (0, 5, 11):
        t, pb = mtr.synth.serving_request(q, n=12, xy=(16, 24), z=(8, 12), theta=[-30.0, 0.0, 30.0])
        assert t == a[q][0] and pb.K == a[q][1].K and pb.vol.X == a[q][1].vol.X and pb.vol.Z == a[q][1].vol.Z
        assert all(np.array_equal(x, y) for x, y in zip(pb.ps, a[q][1].ps))


def test_server_without_device_is_an_error(mtr):
    if mtr.gpu_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(mtr.MtrError) as ei:
        mtr.Server(0)
    assert ei.value.code == -8


def test_server_argument_checks(mtr):
    import ctypes as ct
    out = ct.c_void_p()
    assert mtr.lib.mtr_server_create(0, 7, 1 << 30, 4, ct.byref(out)) == -1  # unknown precision
    assert mtr.lib.mtr_server_create(0, 0, 0, 4, ct.byref(out)) == -1        # empty budget
    assert mtr.lib.mtr_server_create(0, 0, 1 << 30, 0, ct.byref(out)) == -1  # no slots


def test_bench_refuses_more_than_one_gpu():
    # one GPU per box: the bench says so and exits before touching a device
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
