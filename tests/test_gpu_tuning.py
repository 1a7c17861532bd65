"""Launch tuning of the wavefront engine (rtg_set_tuning / rtg_autotune):
tile shape, queue grab and persistent grid only change the order in which
rays are traced, so every frame and every batch must be bit-identical under
every tuning, and the autotuner must install one of the tunings it timed."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(tile_w=4, grab=1, grid_pct=100),
    dict(tile_w=16, grab=2, grid_pct=50),
    dict(tile_w=32, grab=4, grid_pct=75),
    dict(tile_w=8, grab=4, grid_pct=25),
]
DEFAULT = dict(tile_w=8, grab=1, grid_pct=100)


@pytest.fixture
def rt_default(rt_gpu):
    rt_gpu.set_tuning(**DEFAULT)
    yield rt_gpu
    rt_gpu.set_tuning(**DEFAULT)


def _same(a, b, what):
    assert (a["rgb"] == b["rgb"]).all(), f"{what}: rgb differs"
    assert (a["ids"] == b["ids"]).all(), f"{what}: ids differ"
    assert a["stats"].rays_total == b["stats"].rays_total, f"{what}: ray counts differ"


def test_frames_identical_under_every_tuning(rt_default):
    rt = rt_default
    w1 = rt.DeviceScene(rt.HostScene.workload(1))
    w2 = rt.DeviceScene(rt.HostScene.workload(2, width=480, height=270))
    w5 = rt.DeviceScene(rt.HostScene.workload(5, width=1024, height=576, detail=8, n_spheres=64, n_lights=16))
    cases = [
        ("workload 1, ragged 333x211, depth 3", w1, rt.params(333, 211, max_level=3)),
        ("workload 1, 640x480 tile_compat", w1, rt.params(640, 480, max_level=3, tile_compat=1)),
        ("workload 2, 480x270, depth 4", w2, rt.params(480, 270, max_level=4)),
        ("workload 5, 1024x576 4 spp depth 8 (two passes)", w5, rt.params(1024, 576, max_level=8, spp=4, seed=5)),
    ]
    refs = [d.render(p) for _, d, p in cases]
    for t in TUNINGS:
        rt.set_tuning(**t)
        assert rt.get_tuning() == t
        for (what, d, p), ref in zip(cases, refs):
            log(f"case: {what} under tuning {t}")
            _same(d.render(p), ref, f"{what} under {t}")


def test_batch_identical_under_tuning(rt_default):
    rt = rt_default
    hs = [rt.HostScene.workload(3, index=0, detail=20000), rt.HostScene.workload(1)]
    ds = [rt.DeviceScene(h) for h in hs]
    rng = np.random.default_rng(7)
    items = []
    for k in range(12):
        s = k % 2
        w, h = (int(rng.integers(33, 300)), int(rng.integers(17, 200))) if k % 3 == 0 else (128, 64)
        items.append((ds[s], rt.params(w, h, max_level=[2, 3][s], spp=1 + 3 * (k % 4 == 1), seed=k)))
    refs = [d.render(p, tprim=False) for d, p in items]
    for t in TUNINGS[1:3]:
        rt.set_tuning(**t)
        log(f"case: batch of {len(items)} requests (ragged sizes, spp 1/4) under tuning {t}")
        outs, st = rt.render_batch(items, eager=True)
        for k, (o, ref) in enumerate(zip(outs, refs)):
            assert (o["rgb"] == ref["rgb"]).all(), f"request {k} under {t}: rgb differs"
            assert (o["ids"] == ref["ids"]).all(), f"request {k} under {t}: ids differ"
        assert st.rays_total == sum(r["stats"].rays_total for r in refs)


def test_autotune_installs_a_timed_tuning(rt_default):
    rt = rt_default
    d = rt.DeviceScene(rt.HostScene.workload(3, index=1, detail=20000))
    items = [(d, rt.params(w, h, max_level=3, seed=k)) for k, (w, h) in enumerate([(256, 192), (128, 128), (320, 64)])]
    refs = [d.render(p, tprim=False) for _, p in items]
    for what, sel in (("batch of 3", items), ("single frame", items[:1])):
        rt.set_tuning(**DEFAULT)
        log(f"case: autotune over a {what}")
        best, trials = rt.autotune(sel, reps=2)
        log(f"  best {best}; {len(trials)} trials: " +
            ", ".join(f"{t['tile_w']}/{t['grab']}/{t['grid_pct']}={ms:.3f}" for t, ms in trials))
        assert trials and trials[0][0] == DEFAULT
        assert best in [t for t, _ in trials]
        assert min(ms for _, ms in trials) == dict((tuple(t.values()), ms) for t, ms in trials)[tuple(best.values())]
        assert rt.get_tuning() == best
        assert len({tuple(t.values()) for t, _ in trials}) == len(trials)  # each tuning timed once
    outs, _ = rt.render_batch(items, eager=True)
    for o, ref in zip(outs, refs):
        assert (o["rgb"] == ref["rgb"]).all()
