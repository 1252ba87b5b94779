"""The allocation-free forward projection (mtr_op_fp_noalloc) through the C ABI.

It stages its view tables in the library's fixed per-GPU scratch array
instead of allocating, and runs the same kernel as mtr_op_fp: its s and c
must equal mtr_op_fp's bit for bit.  Two calls in a row, the second on a
smaller problem (fewer views, other geometry kind), each checked; the
device outputs are filled with NaN bytes first, so a pixel the launch does
not write fails the comparison.  Size and pointer checks are host-side
rejections.  Device buffers come from the HIP runtime the library itself is
linked against.
"""
import ctypes as ct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip(mtr):
    h = ct.CDLL("libamdhip64.so.7")  # already loaded by libmtr.so: the same runtime
    h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    h.hipFree.argtypes = [ct.c_void_p]
    h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    h.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipMemset, h.hipDeviceSynchronize):
        f.restype = ct.c_int
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, nbytes
        self.p = ct.c_void_p()
        assert hip.hipMalloc(ct.byref(self.p), nbytes) == 0
        assert hip.hipMemset(self.p, 0xFF, nbytes) == 0  # NaN in every float

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.nbytes == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes == self.n
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ct.c_void_p()


CASES = [
    # oblique in-plane axis: general ray plan, 48x40x14 box, 17 views
    dict(name="oblique", X=48, Y=40, Z=14, phi=45.0, kind=0, views=17),
    # second call smaller: tilt axis along Z (row-separable plan), 7 views, ragged detector
    dict(name="axisZ", X=21, Y=17, Z=9, phi=0.0, kind=1, views=7, nu=19, nv=11),
]


def case_inputs(mtr, cs, seed):
    pb = mtr.synth.make_problem(cs["name"], cs["X"], cs["Y"], cs["Z"], 4, [cs["phi"]], kind=cs["kind"],
                                seed=seed, amin=2, amax=5, theta=np.linspace(-60, 60, cs["views"]),
                                nu=cs.get("nu"), nv=cs.get("nv"), noise=0.0)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 2, size=(pb.vol.Z, pb.vol.Y, pb.vol.X)).astype(np.float32)
    return pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], x


def test_noalloc_fp_matches_op_fp_on_two_calls(mtr, hip):
    with mtr.Context(0) as c:
        inputs = [case_inputs(mtr, cs, 40 + q) for q, cs in enumerate(CASES)]
        bufs = []
        try:
            outs = []
            for cs, (prj, vol, xc, yc, x) in zip(CASES, inputs):
                P = prj.X * prj.Y * prj.AngN
                dx, ds, dc = DevBuf(hip, x.nbytes), DevBuf(hip, 4 * P), DevBuf(hip, 4 * P)
                bufs += [dx, ds, dc]
                dx.put(x)
                say(f"noalloc FP call: {cs['name']} {vol.X}x{vol.Y}x{vol.Z}, {prj.AngN} views, "
                    f"detector {prj.X}x{prj.Y}")
                rc = mtr.lib.mtr_op_fp_noalloc(c._h, ct.byref(prj), ct.byref(vol),
                                               xc.ctypes.data_as(ct.POINTER(ct.c_double)),
                                               yc.ctypes.data_as(ct.POINTER(ct.c_double)), dx.p, ds.p, dc.p)
                assert rc == 0, mtr.lib.mtr_last_error()
                shape = (prj.AngN, prj.Y, prj.X)
                outs.append((ds.get(shape), dc.get(shape)))
            for cs, (prj, vol, xc, yc, x), (s, cc) in zip(CASES, inputs, outs):
                s_ref, c_ref = c.op_fp(prj, vol, xc, yc, x)
                say(f"  {cs['name']}: s equal {np.array_equal(s, s_ref)}, c equal {np.array_equal(cc, c_ref)}")
                assert np.isfinite(s).all() and np.isfinite(cc).all()
                assert np.array_equal(s, s_ref) and np.array_equal(cc, c_ref)
                assert float(cc.max()) > 0
        finally:
            for b in bufs:
                b.free()


def test_noalloc_fp_rejects_what_it_cannot_take(mtr, hip):
    prj, vol, xc, yc, x = case_inputs(mtr, CASES[1], 41)
    dp = ct.POINTER(ct.c_double)
    with mtr.Context(0) as c:
        P = prj.X * prj.Y * prj.AngN
        dx, ds, dc = DevBuf(hip, x.nbytes), DevBuf(hip, 4 * P), DevBuf(hip, 4 * P)
        try:
            # more views than the scratch holds
            many = mtr.projection_size(prj.X, prj.Y, 257)  # MTR_NOALLOC_MAX_VIEWS + 1
            big_xc = np.tile(xc[:4], many.AngN)
            big_yc = np.tile(yc[:4], many.AngN)
            rc = mtr.lib.mtr_op_fp_noalloc(c._h, ct.byref(many), ct.byref(vol), big_xc.ctypes.data_as(dp),
                                           big_yc.ctypes.data_as(dp), dx.p, ds.p, dc.p)
            assert rc == -2 and b"at most 256" in mtr.lib.mtr_last_error()
            # a host buffer where device memory is required
            rc = mtr.lib.mtr_op_fp_noalloc(c._h, ct.byref(prj), ct.byref(vol), xc.ctypes.data_as(dp),
                                           yc.ctypes.data_as(dp), ct.c_void_p(x.ctypes.data), ds.p, dc.p)
            assert rc == -1 and b"not device memory" in mtr.lib.mtr_last_error()
            # the context still works after the rejections
            dx.put(x)
            assert mtr.lib.mtr_op_fp_noalloc(c._h, ct.byref(prj), ct.byref(vol), xc.ctypes.data_as(dp),
                                             yc.ctypes.data_as(dp), dx.p, ds.p, dc.p) == 0
            s_ref, _ = c.op_fp(prj, vol, xc, yc, x)
            assert np.array_equal(ds.get((prj.AngN, prj.Y, prj.X)), s_ref)
        finally:
            for b in (dx, ds, dc):
                b.free()
