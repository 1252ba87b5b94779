"""The HBM footprint bound that admission relies on.

`mtr_agent_footprint` is what the request server admits by
(`mtr_server_footprint` adds the per-request tables); it must bound the
device bytes an agent actually holds, and held at most while it was added,
for every kernel family its geometry can select and every storage
precision: the row-separable kernels (with the streaming FP cut into Z
segments, whose partial ray sums are the largest optional buffer), the
column plan and the general plan.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


CASES = [
    # name, in-plane axis, geometry kind, streaming FP segment planes
    ("slab_segments", 0.0, None, "4"),
    ("slab_whole_rays", 90.0, None, "0"),
    ("column_plan", 45.0, None, None),
    ("axis_z", 0.0, 1, None),
]


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("name,phi,kind,seg", CASES, ids=[c[0] for c in CASES])
def test_agent_footprint_bounds_its_device_bytes(mtr, name, phi, kind, seg, prec, monkeypatch):
    if seg is not None:
        monkeypatch.setenv("MTR_FP_STREAM_SEG", seg)
    kw = {} if kind is None else {"kind": kind}
    pb = mtr.synth.make_problem("mem", 48, 40, 24, 6, [phi], seed=12, amin=2, amax=6,
                                theta=np.linspace(-60, 60, 13), **kw)
    with mtr.Context(0, prec) as c:
        c.set_volume(pb.vol)
        bound = c.agent_footprint(pb.prjs[0])
        c.add_agent(pb.prjs[0], pb.xcs[0], pb.ycs[0], pb.ps[0])
        c.bpt(0, out=False)
        c.sirt(0, 1, 0.5, out=False)
        c.sync()
        now, peak = c.agent_device_bytes(0)
        say(f"{name} prec {prec}: kernels {c.agent_kernels(0)}, holds {now}, peak {peak}, bound {bound}")
        if name == "column_plan":
            assert c.agent_kernels(0)[0] == mtr.FPK_COLUMN
        elif name != "axis_z":
            assert c.agent_kernels(0)[0] == mtr.FPK_STREAM
    assert 0 < now <= peak <= bound
