"""Batched rendering (rtg_render_batch): many requests of uneven size, depth
and scene in one wavefront, captured into a hipGraph and replayed.

Each request's frame must be identical to rendering it alone (rtg_render),
and every graph replay identical to the eager launch of the same batch."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu


def _scenes(rt, refraction: bool):
    """Resident scenes: two batch scenes (soup, height field), workload 1
    (plane + spheres) and, with `refraction`, a reduced stress scene whose
    refracting spheres give two child rays per hit (queues too large to bound
    up front: the level loop reads each level's count back)."""
    hs = [rt.HostScene.workload(3, index=0, detail=20000), rt.HostScene.workload(3, index=1, detail=50000),
          rt.HostScene.workload(1)]
    if refraction:
        hs.append(rt.HostScene.workload(5, width=256, height=144, detail=32, n_spheres=24, n_lights=16))
    return hs, [rt.DeviceScene(h) for h in hs]


def _requests(rt, hs, ds, seed, n=24):
    rng = np.random.default_rng(seed)
    items = []
    for k in range(n):
        s = int(rng.integers(0, len(ds)))
        w = int(64 * round(16 ** rng.random())) if k % 5 else int(rng.integers(33, 300))  # also ragged sizes
        h = int(64 * round(12 ** rng.random())) if k % 5 else int(rng.integers(17, 200))
        w, h = min(w, 512), min(h, 384)
        depth = [hs[s].info.max_level, 2, 3, 6][s]
        spp = 4 if s == 3 and k % 2 else 1
        items.append((ds[s], rt.params(w, h, max_level=depth, spp=spp, seed=k, tile_compat=int(k % 7 == 3))))
    return items


@pytest.mark.parametrize("refraction", [False, True], ids=["bounded_queues", "refraction"])
def test_batch_equals_single_renders(rt_gpu, refraction):
    rt = rt_gpu
    hs, ds = _scenes(rt, refraction)
    items = _requests(rt, hs, ds, 1)
    log(f"case: batch of {len(items)} requests over {len(ds)} scenes (refraction={refraction}, "
        f"17x33 .. 512x384, depth 2-6, spp 1/4) vs single renders")
    outs, st = rt.render_batch(items)
    log(f"  batch rays {st.rays_total}, {st.render_ms:.3f} ms")
    total = 0
    for k, ((d, p), o) in enumerate(zip(items, outs)):
        ref = d.render(p, tprim=False)
        assert (o["rgb"] == ref["rgb"]).all(), f"request {k}: rgb differs"
        assert (o["ids"] == ref["ids"]).all(), f"request {k}: ids differ"
        total += ref["stats"].rays_total
    assert st.rays_total == total


def test_batch_graph_replays_match_eager(rt_gpu):
    rt = rt_gpu
    hs, ds = _scenes(rt, refraction=False)
    items = _requests(rt, hs, ds, 2)
    log("case: batch graph capture + 3 replays, eager work between replays")
    eager, st_e = rt.render_batch(items, eager=True)
    first, st0 = rt.render_batch(items)  # captures
    assert st0.graph_replays == 0
    for k, (a, b) in enumerate(zip(first, eager)):
        assert (a["rgb"] == b["rgb"]).all() and (a["ids"] == b["ids"]).all(), f"captured launch, request {k}"
    for rep in range(3):
        ds[2].render(rt.params(200, 100, max_level=3))  # unrelated eager work in between
        again, st = rt.render_batch(items)
        assert st.graph_replays == 1, f"replay {rep}: batch was not replayed from the graph"
        for k, (a, b) in enumerate(zip(again, eager)):
            assert (a["rgb"] == b["rgb"]).all() and (a["ids"] == b["ids"]).all(), f"replay {rep}, request {k}"
        assert st.rays_total == st_e.rays_total
    # reused host buffers, scribbled first: the replay overwrites every pixel
    for o in first:
        o["rgb"].fill(7)
        o["ids"].fill(-7)
    again, st = rt.render_batch(items, out=first)
    assert st.graph_replays == 1 and again[0]["rgb"] is first[0]["rgb"]
    for k, (a, b) in enumerate(zip(again, eager)):
        assert (a["rgb"] == b["rgb"]).all() and (a["ids"] == b["ids"]).all(), f"reused buffers, request {k}"
    # same padded shape, different content: the replay reads the new tables
    items2 = [(d, rt.params(p.width, p.height, max_level=p.max_level, spp=p.spp, seed=p.seed + 100,
                            tile_compat=p.tile_compat)) for d, p in reversed(items)]
    outs2, st2 = rt.render_batch(items2)
    log(f"  reordered batch replayed from the graph: {bool(st2.graph_replays)}")
    for k, ((d, p), o) in enumerate(zip(items2, outs2)):
        ref = d.render(p, tprim=False)
        assert (o["rgb"] == ref["rgb"]).all(), f"reordered request {k}"


def test_batch_single_request_and_empty(rt_gpu):
    rt = rt_gpu
    log("case: batch of one request and of zero requests")
    hs = rt.HostScene.workload(1)
    ds = rt.DeviceScene(hs)
    p = rt.params(640, 480, max_level=3)
    outs, st = rt.render_batch([(ds, p)])
    ref = ds.render(p)
    assert (outs[0]["rgb"] == ref["rgb"]).all() and (outs[0]["ids"] == ref["ids"]).all()
    # a frame into reused (scribbled) host buffers
    for a in (outs[0]["rgb"], outs[0]["ids"]):
        a.fill(3)
    again = ds.render(p, tprim=False, out=outs[0])
    assert again["rgb"] is outs[0]["rgb"] and (again["rgb"] == ref["rgb"]).all() and (again["ids"] == ref["ids"]).all()
    outs, st = rt.render_batch([])
    assert outs == [] and st.rays_total == 0
