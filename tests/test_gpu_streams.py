"""Stream schedules of the wavefront engine: each level's shadow rays run on a
side stream beside the next level, the last level's on the render stream;
RTG_FLAG_ONE_STREAM runs every kernel on the render stream.  The schedule only
changes when kernels run, so frames and batches must be bit-identical under
both, at every depth (depth 0: nothing forked; depth 1: the only forked level
is level 0), and a captured batch graph must replay each schedule it was
captured with."""
from __future__ import annotations

import pytest

from conftest import log

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert (a["rgb"] == b["rgb"]).all(), f"{what}: rgb differs"
    assert (a["ids"] == b["ids"]).all(), f"{what}: ids differ"
    assert a["stats"].rays_total == b["stats"].rays_total, f"{what}: ray counts differ"


def test_frames_identical_on_one_stream(rt_gpu):
    rt = rt_gpu
    w1 = rt.DeviceScene(rt.HostScene.workload(1))
    w2 = rt.DeviceScene(rt.HostScene.workload(2, width=480, height=270))
    cases = [("workload 1, 640x480, depth 3", w1, dict(width=640, height=480, max_level=3)),
             ("workload 1, 640x480, depth 0", w1, dict(width=640, height=480, max_level=0)),
             ("workload 1, 333x211, depth 1", w1, dict(width=333, height=211, max_level=1)),
             ("workload 2, 480x270, depth 4", w2, dict(width=480, height=270, max_level=4)),
             ("workload 2, 1024x576, depth 4 (over 2^18 samples)", w2, dict(width=1024, height=576, max_level=4))]
    for what, d, p in cases:
        log(f"case: {what}, side stream vs one stream")
        _same(d.render(rt.params(**p, flags=rt.RTG_FLAG_ONE_STREAM)), d.render(rt.params(**p)), what)


def test_batches_identical_on_one_stream(rt_gpu):
    rt = rt_gpu
    hs = [rt.HostScene.workload(3, index=0, detail=20000), rt.HostScene.workload(1)]
    ds = [rt.DeviceScene(h) for h in hs]
    shapes = [(128, 64), (96, 80), (200, 120), (64, 64), (128, 128), (160, 96), (64, 128), (256, 64)]
    items = [(ds[k % 2], dict(width=w, height=h, max_level=[2, 3][k % 2], seed=k)) for k, (w, h) in enumerate(shapes)]
    refs = [d.render(rt.params(**p), tprim=False) for d, p in items]
    for flags, name in ((0, "side stream"), (rt.RTG_FLAG_ONE_STREAM, "one stream")):
        batch = [(d, rt.params(**p, flags=flags)) for d, p in items]
        log(f"case: batch of {len(batch)} on the {name}: eager, then captured and replayed")
        runs = [rt.render_batch(batch, eager=True)] + [rt.render_batch(batch) for _ in range(4)]
        assert runs[-1][1].graph_replays == 1, f"{name}: the batch was not replayed from a graph"
        for r, (outs, st) in enumerate(runs):
            for k, (o, ref) in enumerate(zip(outs, refs)):
                assert (o["rgb"] == ref["rgb"]).all(), f"{name}, run {r}, request {k}: rgb differs"
                assert (o["ids"] == ref["ids"]).all(), f"{name}, run {r}, request {k}: ids differ"
            assert st.rays_total == sum(x["stats"].rays_total for x in refs)
