"""Serving loop (rtg_serve): Poisson arrivals over resident scenes,
continuous batching over concurrent slots, against the plain renderer.

Each request's frame must be the same when it is served in the mix as when
it is rendered alone (rtg_render): bit-identical RGB."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu


def _setup(rt, n=48, rate=4000.0):
    hs = [rt.HostScene.workload(3, index=i, detail=[20000, 50000, 8000, 30000][i]) for i in range(4)]
    hs.append(rt.HostScene.workload(1))
    ds = [rt.DeviceScene(h) for h in hs]
    items = []
    for (s, w, h, t) in rt.gen_requests(n, rate, seed=4):
        s = s % len(ds)
        w, h = min(w, 320), min(h, 256)
        items.append((ds[s], rt.params(w, h, max_level=hs[s].info.max_level), t))
    return ds, items


@pytest.mark.parametrize("mode", [0, 1], ids=["batched", "plain"])
def test_serve_same_alone_and_in_mix(rt_gpu, mode):
    rt = rt_gpu
    ds, items = _setup(rt)
    log(f"case: serve {len(items)} Poisson requests at 4000/s over 5 scenes, mode {'batched plain'.split()[mode]}")
    outs, done, st = rt.serve(items, mode=mode, slots=2)
    log(f"  p50 {st.p50_ms:.2f} ms  p99 {st.p99_ms:.2f} ms  {st.mrays_per_s:.1f} Mrays/s  batches {st.batches}  "
        f"max in flight {st.max_in_flight}")
    rays = 0
    for k, ((d, p, t), rgb, (ds_, b)) in enumerate(zip(items, outs, done)):
        ref = d.render(p, ids=False, tprim=False)
        assert (rgb == ref["rgb"]).all(), f"request {k} (batch {b}) differs from its render alone"
        assert ds_ >= t and b >= 0
        rays += ref["stats"].rays_total
    assert st.rays_total == rays
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
    if mode == 0:
        assert st.batches < len(items)  # arrivals overlapped the GPU work: requests were batched together
    else:
        assert st.batches == len(items)


def test_serve_budget_splits_batches(rt_gpu):
    rt = rt_gpu
    ds, items = _setup(rt, n=24, rate=0.0)  # all arrive at once
    log("case: serve 24 simultaneous requests under a 64 MiB queue budget")
    outs, done, st = rt.serve(items, slots=2, hbm_budget=64 << 20)
    assert st.batches > 2  # the budget forced several batches
    for (d, p, t), rgb in zip(items, outs):
        assert (rgb == d.render(p, ids=False, tprim=False)["rgb"]).all()
