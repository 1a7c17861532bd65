"""bench.py on the CPU: the one-GPU contract and the algorithmic-traffic model
that the roofline figure divides by (per traced ray of each kind; shadow rays
that cannot change the colour are not traced and cost 0)."""
from __future__ import annotations

import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def _traffic(**kw):
    base = dict(rays=0, node_visits=0, tri_tests=0, sph_tests=0, shadow_rays=0, shadow_node_visits=0,
                shadow_tri_tests=0, shadow_sph_tests=0, shadow_skipped=0)
    base.update(kw)
    return SimpleNamespace(**base)


def test_one_shadow_ray_and_one_closest_ray():
    # one closest-hit ray: 3 nodes, 2 triangles; one traced shadow ray: 5 nodes,
    # 1 triangle, 1 sphere; one skipped shadow ray
    tr = _traffic(rays=2, node_visits=8, tri_tests=3, sph_tests=1, shadow_rays=1, shadow_node_visits=5,
                  shadow_tri_tests=1, shadow_sph_tests=1, shadow_skipped=1)
    m = bench.traffic_model(tr)
    closest = 32 + 32 + 3 * 64 + 2 * 48
    shadow = 32 + 32 + 5 * 64 + 48 + 16
    assert m["per_closest_ray"] == pytest.approx(closest)
    assert m["per_shadow_ray"] == pytest.approx(shadow)
    # three reference-semantic rays, the skipped one at 0 bytes
    assert m["per_reference_ray"] == pytest.approx((closest + shadow) / 3)
    assert m["shadow_skipped_frac"] == pytest.approx(0.5)


def test_traffic_scales_with_rays():
    one = bench.traffic_model(_traffic(rays=1, node_visits=10, tri_tests=4))
    many = bench.traffic_model(_traffic(rays=1000, node_visits=10000, tri_tests=4000))
    assert one["per_closest_ray"] == pytest.approx(many["per_closest_ray"])
    assert one["per_shadow_ray"] == 0.0  # no shadow rays: nothing divided by zero


def test_pixel_error_statistics_count_only_the_sampled_tiles():
    import numpy as np
    # a 130 x 100 frame: 3 x 2 tiles of 64 (the last column and row partial);
    # tiles 0 and 5 (the bottom-right corner, 2 x 36 pixels) are sampled
    W, H = 130, 100
    g = {"ids": np.zeros((H, W, 2), np.int32), "rgb": np.full((H, W, 3), 100, np.uint8)}
    o = {"ids": g["ids"].copy(), "rgb": g["rgb"].copy()}
    o["ids"][3, 5, 1] = 7        # tile 0: one id mismatch
    o["rgb"][10, 10] = (103, 100, 100)  # tile 0: off by 3 in one channel
    o["rgb"][99, 129] = (101, 99, 100)  # tile 5: off by 1 in two channels
    o["ids"][70, 10] = (1, 1)    # tile 3: not sampled
    o["rgb"][70, 10] = (0, 0, 0)
    c = bench.fp64_counts(g, o, W, H, [0, 5])
    px = 64 * 64 + 2 * 36
    assert c == [px, 1, 5.0, 3, 1]
    st = bench.fp64_stats(c)
    assert st["pixels"] == px
    assert st["ids_mismatch"] == round(1 / px, 6)
    assert st["rgb_mae"] == round(5 / (3 * px), 4)
    assert st["rgb_max"] == 3 and st["rgb_over_1"] == round(1 / px, 6)
    # the one-frame statistics are these counts
    assert bench.fp64_error(g, o, W, H, [0, 5]) == st
    # counts of two frames add (the maximum takes the larger)
    two = [a + b if k != 3 else max(a, b) for k, (a, b) in enumerate(zip(c, c))]
    assert bench.fp64_stats(two)["ids_mismatch"] == st["ids_mismatch"] and bench.fp64_stats(two)["pixels"] == 2 * px


def test_request_sample_takes_the_first_request_of_each_scene():
    # scenes 0..3; scene 2 has no request; request 4 is scene 0's second
    reqs = [(1, 64, 64, 0.0), (0, 130, 100, 0.0), (3, 1024, 768, 0.0), (1, 200, 200, 0.0), (0, 64, 64, 0.0)]
    s = bench.sample_requests([None] * 4, reqs, tiles_per=4, seed=7)
    assert [(sc, w, h) for sc, w, h, _ in s] == [(0, 130, 100), (1, 64, 64), (3, 1024, 768)]
    for sc, w, h, tiles in s:
        n = ((w + 63) // 64) * ((h + 63) // 64)
        assert len(tiles) == min(4, n) and len(set(tiles.tolist())) == len(tiles)
        assert all(0 <= t < n for t in tiles) and list(tiles) == sorted(tiles)
    # seeded: the same sample every time
    s2 = bench.sample_requests([None] * 4, reqs, tiles_per=4, seed=7)
    assert all((a[3] == b[3]).all() for a, b in zip(s, s2))
