"""The batch and serving workloads at their full size, against the REF32 oracle.

Batch: the 256 requests (64x64 .. 1024x768) over the 8 resident scenes
(triangle soups and height fields of 1k .. 618k triangles, 1-8 lights, depth
2-6) that the benchmark renders.  Every request of the batch must be
bit-identical to its render alone, and the render alone must meet the parity
bar against the oracle (RayTracer.cpp:450-596) on a seeded 64x64 tile of one
request per scene.

Serving: the Poisson request mix at full sizes through the continuous-batching
loop; every served frame bit-identical to its render alone, plus an oracle
spot check of served frames.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import compare_frames, log, tile_mask

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_scenes(rt_gpu):
    rt = rt_gpu
    log("setup: the 8 resident batch scenes at full size")
    hs = [rt.HostScene.workload(3, index=i) for i in range(8)]
    ds = [rt.DeviceScene(h) for h in hs]
    log("  triangles " + ", ".join(str(int(h.info.n_tris)) for h in hs))
    return hs, ds


def _oracle_spot_check(rt, hs, items, picks, seed):
    """REF32 oracle on one seeded 64x64 tile of each picked request, against the
    request's render alone (threads: the box's CPU share)."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    for k in picks:
        d, p, s = items[k]
        W, H = p.width, p.height
        tx, ty = (W + 63) // 64, (H + 63) // 64
        tile = int(rng.integers(0, tx * ty))
        log(f"  oracle: request {k} (scene {s}, {int(hs[s].info.n_tris)} triangles, {W}x{H}, depth {p.max_level}),"
            f" tile {tile}")
        o = po.OracleScene(hs[s].desc_ptr).render(p, po.ORACLE_REF32, threads=16, tiles=[tile])
        g = d.render(p)
        st = compare_frames(g, o, region=tile_mask(W, H, [tile]))
        log(f"    {st}")
        assert st["ids_mismatch_outside_band"] == 0, (k, st)
        assert st["ids_match"] >= 0.9999, (k, st)
        assert st["rgb_over_1"] == 0, (k, st)


def _one_per_scene(scenes_of):
    picks = {}
    for k, s in enumerate(scenes_of):
        picks.setdefault(s, k)
    return [picks[s] for s in sorted(picks)]


def test_config3_batch_at_full_size(rt_gpu, batch_scenes):
    rt = rt_gpu
    hs, ds = batch_scenes
    reqs = rt.gen_requests(256, 0.0, seed=3)
    items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level), s) for s, w, h, _ in reqs]
    px = sum(p.width * p.height for _, p, _ in items)
    log(f"case: batch of 256 requests, {px} pixels, over the 8 scenes vs renders alone and the oracle")
    outs, st = rt.render_batch([(d, p) for d, p, _ in items])
    log(f"  batch rays {st.rays_total}, {st.render_ms:.2f} ms device")
    total = 0
    for k, ((d, p, s), o) in enumerate(zip(items, outs)):
        ref = d.render(p, tprim=False)
        assert (o["rgb"] == ref["rgb"]).all(), f"request {k}: rgb differs from its render alone"
        assert (o["ids"] == ref["ids"]).all(), f"request {k}: ids differ from its render alone"
        total += ref["stats"].rays_total
    assert st.rays_total == total
    picks = _one_per_scene([s for _, _, s in items])
    assert len(picks) == 8
    _oracle_spot_check(rt, hs, items, picks, seed=31)


def test_config4_serving_mix_at_full_size(rt_gpu, batch_scenes):
    rt = rt_gpu
    hs, ds = batch_scenes
    reqs = rt.gen_requests(200, 1000.0, seed=4)
    items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level), t) for s, w, h, t in reqs]
    log("case: serve 200 Poisson requests at 1000/s, full sizes, over the 8 scenes (batched, 2 slots)")
    outs, done, st = rt.serve(items, slots=2)
    log(f"  p50 {st.p50_ms:.2f} ms  p99 {st.p99_ms:.2f} ms  {st.req_per_s:.0f} req/s  batches {st.batches}  "
        f"max in flight {st.max_in_flight}")
    rays = 0
    for k, ((d, p, t), rgb, (ds_, b)) in enumerate(zip(items, outs, done)):
        ref = d.render(p, ids=False, tprim=False)
        assert (rgb == ref["rgb"]).all(), f"request {k} (batch {b}) differs from its render alone"
        assert ds_ >= t and b >= 0
        rays += ref["stats"].rays_total
    assert st.rays_total == rays
    assert st.batches < len(items)
    # served frames against the oracle: the served RGB equals the render alone
    # (above); the render alone meets the bar on a tile of a few requests
    scenes_of = [s for s, _, _, _ in reqs]
    picks = [k for k in _one_per_scene(scenes_of) if scenes_of[k] != 0][:4]
    _oracle_spot_check(rt, hs, [(d, p, s) for (d, p, _), s in zip(items, scenes_of)], picks, seed=41)
