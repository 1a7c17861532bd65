"""GPU parity: librtg's HIP kernels against the REF32 oracle, through the C ABI.

Bar (per workload): primary (object, triangle) ids equal on >= 99.99% of
pixels, every mismatch inside the oracle's near-tie / cut-edge band; RGB
within 1/255 per channel wherever the primary ids match; reference-semantic
ray counts equal.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import compare_frames, log, sample_tiles, tile_mask

pytestmark = pytest.mark.gpu


def _oracle(hs, p, tiles=None):
    from oracle import pyoracle as po
    return po.OracleScene(hs.desc_ptr).render(p, po.ORACLE_REF32, tiles=tiles)


def _check(stats, min_ids=0.9999):
    assert stats["ids_mismatch_outside_band"] == 0, stats
    assert stats["ids_match"] >= min_ids, stats
    assert stats["rgb_over_1"] == 0, stats


ENGINES = [1, 2]  # megakernel, wavefront


@pytest.mark.parametrize("engine", ENGINES)
def test_workload1_full_frame(rt_gpu, engine):
    rt = rt_gpu
    log(f"case: workload 1, 640x480 full frame, engine {engine}")
    hs = rt.HostScene.workload(1)
    p = rt.params(640, 480, max_level=hs.info.max_level, engine=engine)
    ds = rt.DeviceScene(hs)
    g = ds.render(p)
    o = _oracle(hs, p)
    s = compare_frames(g, o)
    log(f"  {s}")
    _check(s)
    st, ost = g["stats"], o["stats"]
    assert (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract) == (
        ost.rays_primary, ost.rays_shadow, ost.rays_reflect, ost.rays_refract)


def test_workload1_tile_compat(rt_gpu):
    rt = rt_gpu
    log("case: workload 1, RayTracer::start tile semantics (whole 64x64 tiles only)")
    hs = rt.HostScene.workload(1)
    p = rt.params(640, 480, max_level=3, tile_compat=1)
    g = rt.DeviceScene(hs).render(p)
    o = _oracle(hs, p)
    assert (g["rgb"][448:] == 127).all()  # 480 rows -> 7 whole tiles of 64 rows
    assert (g["rgb"][:448] == o["rgb"][:448]).mean() > 0.999
    s = compare_frames(g, o, region=tile_mask(640, 480, range(70)))
    _check(s)


@pytest.mark.parametrize("engine", ENGINES)
def test_workload2_reduced_res_full_mesh(rt_gpu, engine):
    rt = rt_gpu
    log(f"case: workload 2, 480x270, 200k-triangle icosphere, 64 spheres, engine {engine}")
    hs = rt.HostScene.workload(2, width=480, height=270)
    p = rt.params(480, 270, max_level=4, engine=engine)
    ds = rt.DeviceScene(hs)
    log(f"  bvh build {ds.build_ms:.2f} ms, scene bytes {ds.nbytes}")
    g = ds.render(p)
    o = _oracle(hs, p)
    s = compare_frames(g, o)
    log(f"  {s}")
    _check(s)


@pytest.mark.parametrize("engine", ENGINES)
def test_workload2_1080p_sampled_tiles(rt_gpu, engine):
    rt = rt_gpu
    log(f"case: workload 2, 1920x1080, oracle on a 1/16 tile sample, engine {engine}")
    hs = rt.HostScene.workload(2)
    p = rt.params(1920, 1080, max_level=4, engine=engine)
    g = rt.DeviceScene(hs).render(p)
    tiles = sample_tiles(1920, 1080, 16)
    o = _oracle(hs, p, tiles=tiles)
    s = compare_frames(g, o, region=tile_mask(1920, 1080, tiles))
    log(f"  {s}  rays {g['stats'].rays_total}  render {g['stats'].render_ms:.3f} ms")
    _check(s)


def test_determinism_and_scene_reuse(rt_gpu):
    rt = rt_gpu
    log("case: two renders of one resident scene are identical")
    hs = rt.HostScene.workload(2, width=320, height=180, detail=40)
    ds = rt.DeviceScene(hs)
    p = rt.params(320, 180, max_level=4)
    a, b = ds.render(p), ds.render(p)
    assert (a["rgb"] == b["rgb"]).all() and (a["ids"] == b["ids"]).all() and (a["t"] == b["t"]).all()


def test_async_and_cancel(rt_gpu):
    rt = rt_gpu
    from raytrace_amd import _capi
    lib = _capi.load()
    log("case: rtg_render_async + rtg_poll + rtg_job_wait, then a job cancelled before it starts")
    hs = rt.HostScene.workload(1)
    ds = rt.DeviceScene(hs)
    p = rt.params(640, 480, max_level=3)
    rgb = np.zeros((480, 640, 3), np.uint8)
    job = C.c_void_p()
    _capi.check(lib.rtg_render_async(ds.h, C.byref(p), rgb.ctypes.data_as(C.c_void_p), None, None, C.byref(job)),
                "rtg_render_async")
    st = _capi.rtg_stats()
    assert lib.rtg_job_wait(job, C.byref(st)) == 0
    lib.rtg_job_free(job)
    ref = ds.render(p)["rgb"]
    assert (rgb == ref).all()
    # the next frame of the arena is not affected by a job cancelled while it
    # waited for another job: it renders nothing, every pixel stays 127
    import time
    hl = rt.HostScene.workload(5, width=1024, height=576, detail=8, n_spheres=64)
    dl = rt.DeviceScene(hl)
    pl = rt.params(1024, 576, max_level=8, spp=4)
    t0 = time.perf_counter()
    full = dl.render(pl, ids=False, tprim=False)["rgb"]
    frame_s = time.perf_counter() - t0
    log(f"  long frame {frame_s * 1e3:.1f} ms")
    assert frame_s > 0.03
    rgb_a = np.zeros_like(full)
    job_a, job_b = C.c_void_p(), C.c_void_p()
    _capi.check(lib.rtg_render_async(dl.h, C.byref(pl), rgb_a.ctypes.data_as(C.c_void_p), None, None,
                                     C.byref(job_a)), "rtg_render_async")
    time.sleep(frame_s / 4)  # job a holds the arena
    rgb2 = np.zeros_like(rgb)
    _capi.check(lib.rtg_render_async(ds.h, C.byref(p), rgb2.ctypes.data_as(C.c_void_p), None, None,
                                     C.byref(job_b)), "rtg_render_async")
    lib.rtg_cancel(job_b)
    rc_a, rc_b = lib.rtg_job_wait(job_a, None), lib.rtg_job_wait(job_b, None)
    lib.rtg_job_free(job_a)
    lib.rtg_job_free(job_b)
    assert rc_a == 0 and (rgb_a == full).all()
    assert rc_b == -4 and (rgb2 == 127).all()
    assert (ds.render(p)["rgb"] == ref).all()


def _cancel_mid_frame(rt, ds, p, full, frame_s):
    """Start an asynchronous render, cancel it a third of the way through while
    rtg_poll still says pending, and return (status, rgb)."""
    import time
    from raytrace_amd import _capi
    lib = _capi.load()
    rgb = np.zeros_like(full)
    job = C.c_void_p()
    _capi.check(lib.rtg_render_async(ds.h, C.byref(p), rgb.ctypes.data_as(C.c_void_p), None, None, C.byref(job)),
                "rtg_render_async")
    time.sleep(frame_s / 4)
    pending = lib.rtg_poll(job) == 1
    lib.rtg_cancel(job)
    rc = lib.rtg_job_wait(job, None)
    lib.rtg_job_free(job)
    return pending, rc, rgb


@pytest.mark.parametrize("engine", ENGINES)
def test_cancel_mid_frame(rt_gpu, engine):
    """rtg_cancel / RayTracer::stop (RayTracer.cpp:29-33, 698-701) on a frame in
    flight: the job reports RTG_E_CANCELLED, pixels rendered before the cancel
    equal the full render, and the rest keep the value 127."""
    import time
    rt = rt_gpu
    log(f"case: cancel mid-frame, workload 5 at 1024x576, 4 spp, engine {engine}")
    hs = rt.HostScene.workload(5, width=1024, height=576, detail=8, n_spheres=64)
    ds = rt.DeviceScene(hs)
    ml = 8 if engine == 2 else 3
    p = rt.params(1024, 576, max_level=ml, spp=4, engine=engine)
    full = ds.render(p, ids=False, tprim=False)["rgb"]  # warm
    t0 = time.perf_counter()
    full = ds.render(p, ids=False, tprim=False)["rgb"]
    frame_s = time.perf_counter() - t0
    log(f"  full frame {frame_s * 1e3:.1f} ms")
    assert frame_s > 0.03, "frame too short to cancel part-way"
    pending, rc, rgb = _cancel_mid_frame(rt, ds, p, full, frame_s)
    assert pending, "the job finished before the cancel"
    assert rc == -4, rc
    same = (rgb == full).all(-1)
    blank = (rgb == 127).all(-1) & (full != 127).any(-1)
    log(f"  rendered before the cancel: {same.mean():.3f}, untouched: {blank.mean():.3f}")
    assert (same | blank).all()
    assert blank.mean() > 0.001  # the cancel stopped the frame part-way
    if engine == 2:  # a cancellable wavefront frame runs in passes of whole pixel-block rows
        rows_done = same.all(1) | (full == 127).all(-1).all(1)
        rows_blank = blank.all(1)
        assert (rows_done | rows_blank).all()


def test_raytracer_stop(rt_gpu):
    """RayTracer::stop() on a frame in flight, through the drop-in class."""
    import time
    rt = rt_gpu
    log("case: RayTracer::start then RayTracer::stop, workload 5 at 2048x1152")
    hs = rt.HostScene.workload(5, width=2048, height=1152, detail=8, n_spheres=64)
    t0 = time.perf_counter()
    full, use_time = hs.raytrace(0x80, max_level=8)
    log(f"  full frame {time.perf_counter() - t0:.2f} s (useTime {use_time:.3f} s)")
    full, use_time = hs.raytrace(0x80, max_level=8)  # warm: the scene is resident
    out, rc, finished = hs.raytrace_stopped(0x80, 8, use_time / 4)
    assert not finished, "the frame finished before the stop"
    assert rc == -4, rc
    W, H = 2048, 1152
    a = out[:H * W * 3].reshape(H, W, 3)
    f = full[:H * W * 3].reshape(H, W, 3)
    same, blank = (a == f).all(-1), (a == 127).all(-1) & (f != 127).any(-1)
    log(f"  rendered before the stop: {same.mean():.3f}, untouched: {blank.mean():.3f}")
    assert (same | blank).all() and blank.mean() > 0.001


def test_raytracer_drop_in(rt_gpu):
    """RayTracer::start / isFinish / output through the drop-in C++ class."""
    rt = rt_gpu
    log("case: RayTracer::start(MY_MODEL_RAYTRACE) on workload 1")
    hs = rt.HostScene.workload(1)
    buf, use_time = hs.raytrace(0x80, 3)
    img = buf[:640 * 480 * 3].reshape(480, 640, 3)
    o = _oracle(hs, rt.params(640, 480, max_level=3, tile_compat=1))
    assert (img[448:] == 127).all()
    d = np.abs(img[:448].astype(int) - o["rgb"][:448].astype(int)).max()
    assert d <= 1
    assert (buf[640 * 480 * 3:] == 127).all()
    assert use_time > 0
    # re-rendering an unchanged scene keeps it resident: one upload for three frames
    buf3, _, uploads = hs.raytrace(0x80, 3, frames=3, with_uploads=True)
    assert uploads == 1
    assert (buf3 == buf).all()


@pytest.mark.parametrize("case", ["w2_1080p", "w5_refraction_spp4", "w1_tile_compat", "w3_heightfield"])
def test_engines_bit_identical(rt_gpu, case):
    """Megakernel and wavefront compose every ray-tree node with the same
    operations: frames, ids and ray counts must be identical."""
    rt = rt_gpu
    log(f"case: megakernel vs wavefront, {case}")
    if case == "w2_1080p":
        hs, p = rt.HostScene.workload(2), dict(width=1920, height=1080, max_level=4)
    elif case == "w5_refraction_spp4":
        hs = rt.HostScene.workload(5, width=240, height=136, detail=8, n_spheres=64, n_lights=16)
        p = dict(width=240, height=136, max_level=8, spp=4, seed=5)
    elif case == "w1_tile_compat":
        hs, p = rt.HostScene.workload(1, width=650, height=500), dict(width=650, height=500, max_level=3, tile_compat=1)
    else:
        hs = rt.HostScene.workload(3, index=1, width=320, height=256, detail=200000)
        p = dict(width=320, height=256, max_level=hs.info.max_level)
    ds = rt.DeviceScene(hs)
    a = ds.render(rt.params(engine=1, **p))
    b = ds.render(rt.params(engine=2, **p))
    assert (a["rgb"] == b["rgb"]).all()
    assert (a["ids"] == b["ids"]).all() and (a["t"] == b["t"]).all()
    sa, sb = a["stats"], b["stats"]
    assert (sa.rays_primary, sa.rays_shadow, sa.rays_reflect, sa.rays_refract) == (
        sb.rays_primary, sb.rays_shadow, sb.rays_reflect, sb.rays_refract)
    log(f"  rays {sa.rays_total}  mega {sa.render_ms:.3f} ms  wave {sb.render_ms:.3f} ms")


def test_workload5_reduced_vs_oracle(rt_gpu):
    """Stress features (refraction through sphere meshes, 16 lights, 4 jittered
    samples, depth 8) at reduced size against the oracle on a tile sample."""
    rt = rt_gpu
    log("case: workload 5 reduced (960x540, terrain /4, 256 sphere meshes), oracle on 3 tiles")
    hs = rt.HostScene.workload(5, width=960, height=540, detail=4)
    p = rt.params(960, 540, max_level=8, spp=4, seed=5, engine=2)
    g = rt.DeviceScene(hs).render(p)
    log(f"  gpu rays {g['stats'].rays_total}, {g['stats'].render_ms:.2f} ms; oracle follows")
    tiles = np.array([4 * 15 + 6, 4 * 15 + 7, 5 * 15 + 7], np.int32)  # centre of the frame (15 tiles per row)
    o = _oracle(hs, p, tiles=tiles)
    s = compare_frames(g, o, region=tile_mask(960, 540, tiles))
    log(f"  {s}  rays {g['stats'].rays_total}")
    _check(s)


def test_fp64_retest_on_device(rt_gpu):
    """RTG_FLAG_FP64_RETEST on the GPU: identical to the device code run on the
    CPU with the flag, and different from the reference algorithm only inside
    the oracle's near-tie band."""
    from conftest import emul_render
    rt = rt_gpu
    log("case: workload 2 at 480x270 with the FP64 near-tie re-test")
    hs = rt.HostScene.workload(2, width=480, height=270)
    pf = rt.params(480, 270, max_level=4, flags=rt.RTG_FLAG_FP64_RETEST)
    g = rt.DeviceScene(hs).render(pf)
    e = emul_render(hs.desc_ptr, pf)
    assert (g["rgb"] == e["rgb"]).all() and (g["ids"] == e["ids"]).all() and (g["t"] == e["t"]).all()
    assert g["stats"].retests_fp64 == e["retests"] > 0
    s = compare_frames(g, _oracle(hs, rt.params(480, 270, max_level=4)))
    log(f"  {s}  re-tests {g['stats'].retests_fp64}")
    assert s["ids_mismatch_outside_band"] == 0 and s["rgb_over_1"] == 0


@pytest.mark.parametrize("workload", [1, 2, 3, 5])
def test_fp64_error_statistics(rt_gpu, workload):
    """Pixel-error statistics of the device frames (reference FP32 semantics,
    and with the FP64 near-tie re-test) against the FP64 oracle."""
    from conftest import fp64_stats
    from oracle import pyoracle as po
    rt = rt_gpu
    if workload == 1:
        hs, pr = rt.HostScene.workload(1), dict(width=640, height=480, max_level=3)
    elif workload == 2:
        hs, pr = rt.HostScene.workload(2, width=480, height=270), dict(width=480, height=270, max_level=4)
    elif workload == 3:
        hs = rt.HostScene.workload(3, index=1, width=320, height=256)
        pr = dict(width=320, height=256, max_level=hs.info.max_level)
    else:
        hs = rt.HostScene.workload(5, width=320, height=180, detail=8, n_spheres=64, n_lights=16)
        pr = dict(width=320, height=180, max_level=8, spp=4, seed=5)
    ds = rt.DeviceScene(hs)
    g = ds.render(rt.params(**pr))
    gf = ds.render(rt.params(**pr, flags=rt.RTG_FLAG_FP64_RETEST))
    o64 = po.OracleScene(hs.desc_ptr).render(rt.params(**pr), po.ORACLE_FP64)
    a, b = fp64_stats(g, o64), fp64_stats(gf, o64)
    log(f"case: workload {workload} vs FP64 oracle: fp32 {a}; fp32+fp64 re-test {b}, re-tests {gf['stats'].retests_fp64}")
    assert a["ids_mismatch"] < 0.01 and b["ids_mismatch"] <= a["ids_mismatch"] + 1e-4


def test_branching_frame_in_passes(rt_gpu):
    """A refraction frame larger than one pass (2^21 level-0 samples) renders
    in passes of block rows; pixels on both sides of the pass boundary match
    the device code run on the CPU and the reference algorithm."""
    from conftest import emul_render
    rt = rt_gpu
    W, H = 1024, 576  # 4 spp: 2,359,296 samples, two passes (block rows 0-63 | 64-71)
    log(f"case: workload 5 (terrain /8, 64 sphere meshes) at {W}x{H} 4 spp, two passes")
    hs = rt.HostScene.workload(5, width=W, height=H, detail=8, n_spheres=64, n_lights=16)
    p = rt.params(W, H, max_level=8, spp=4, seed=5)
    g = rt.DeviceScene(hs).render(p)
    tx = (W + 63) // 64
    tiles = np.array([7 * tx + 7, 8 * tx + 8, 8 * tx + 3], np.int32)  # rows 448-511 and 512-575 (pass boundary at 512)
    e = emul_render(hs.desc_ptr, p, tiles=tiles)
    m = tile_mask(W, H, tiles)
    assert (g["rgb"][m] == e["rgb"][m]).all() and (g["ids"][m] == e["ids"][m]).all()
    o = _oracle(hs, p, tiles=tiles)
    s = compare_frames(g, o, region=m)
    log(f"  {s}  rays {g['stats'].rays_total}  {g['stats'].render_ms:.1f} ms")
    _check(s)


@pytest.mark.parametrize("mode", [0x1, 0x2, 0x3, 0x4, 0x5, 0x6, 0x7])
def test_debug_modes_on_device(rt_gpu, mode):
    """RayTracer::start(type) debug modes on the GPU against the reference
    algorithm (and bit for bit against the device code run on the CPU)."""
    from conftest import emul_render
    rt = rt_gpu
    log(f"case: render mode {mode:#x} on workload 1 (and on a refraction scene for the reflect test)")
    cases = [(rt.HostScene.workload(1), dict(width=640, height=480, max_level=3))]
    if mode == 0x7:
        cases.append((rt.HostScene.workload(5, width=256, height=144, detail=32, n_spheres=24, n_lights=16),
                      dict(width=256, height=144, max_level=6)))
    for hs, pr in cases:
        ds = rt.DeviceScene(hs)
        for engine in ([1, 2] if mode == 0x7 else [0]):
            p = rt.params(mode=mode, engine=engine, **pr)
            g = ds.render(p)
            o = _oracle(hs, p)
            s = compare_frames(g, o)
            _check(s)
            e = emul_render(hs.desc_ptr, p)
            assert (g["rgb"] == e["rgb"]).all() and (g["ids"] == e["ids"]).all()
            st, ost = g["stats"], o["stats"]
            assert (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract) == (
                ost.rays_primary, ost.rays_shadow, ost.rays_reflect, ost.rays_refract)
    # through the drop-in class: whole 64x64 tiles only, the rest stays 127
    buf, _ = cases[0][0].raytrace(mode, 3)
    img = buf[:640 * 480 * 3].reshape(480, 640, 3)
    assert (img[448:] == 127).all()


def test_scene_update_paths(rt_gpu):
    """rtg_scene_update (the reference's Scene mutations, Scene.cpp:157-301):
    material, light and camera changes re-upload the shading tables only (the
    BVH stays: same build time); moving an object rebuilds.  Each updated
    scene renders exactly as a freshly created one, including a material that
    turns refraction on (two child rays per hit: the queue bounds follow)."""
    rt = rt_gpu
    log("case: scene updates on workload 1: material (adds refraction), light, object move")
    hs = rt.HostScene.workload(1)
    ds = rt.DeviceScene(hs)
    p = rt.params(640, 480, max_level=4)
    b0 = ds.build_ms

    def same_as_fresh():
        g, f = ds.render(p), rt.DeviceScene(hs).render(p)
        assert (g["rgb"] == f["rgb"]).all() and (g["ids"] == f["ids"]).all()
        _check(compare_frames(g, _oracle(hs, p)))

    hs.set_material(1, (0.1, 0.1, 0.1), (0.3, 0.4, 0.5), (0.5, 0.5, 0.5), 20.0, reflect=0.2, refract=0.6, rfr=1.4)
    ds.update(hs)
    assert ds.build_ms == b0  # shading tables only
    same_as_fresh()
    hs.set_light_position(0, (2.0, 9.0, 3.0, 1.0))
    ds.update(hs)
    assert ds.build_ms == b0
    same_as_fresh()
    hs.move_pos(2, 2, (0.5, 0.25, -0.5))  # MovePos(MY_MODEL_OBJECT, 2): same primitives, moved
    ds.update(hs)
    st = ds.stats()
    assert st["refits"] == 1 and st["rebuilds"] == 0 and st["shading_updates"] == 2
    same_as_fresh()


def test_scene_refit_on_move(rt_gpu):
    """MovePos (Scene.cpp:190-230) keeps the primitive set, so rtg_scene_update
    refits the BVH over its kept topology instead of rebuilding it; every
    refitted scene renders exactly as a freshly built one (and as the
    reference algorithm).  Switch changes the primitive set: rebuilt.
    rtg_scene_rebuild always rebuilds."""
    rt = rt_gpu
    log("case: workload 2 (480x270): four MovePos refits, a Switch rebuild, a forced rebuild")
    hs = rt.HostScene.workload(2, width=480, height=270)
    ds = rt.DeviceScene(hs)
    p = rt.params(480, 270, max_level=4)

    def same_as_fresh(what):
        g, f = ds.render(p), rt.DeviceScene(hs).render(p)
        assert (g["rgb"] == f["rgb"]).all() and (g["ids"] == f["ids"]).all(), what
        assert g["stats"].rays_total == f["stats"].rays_total, what
        return g

    moves = [(0, (0.3, 0.0, 0.0)), (3, (0.0, 1.5, 0.0)), (0, (-1.0, 0.5, 0.2)), (10, (2.0, -0.5, 1.0))]
    for k, (obj, v) in enumerate(moves):
        hs.move_pos(2, obj, v)
        ds.update(hs)
        st = ds.stats()
        log(f"  move {k}: object {obj} by {v}: {st}")
        assert st["refits"] == k + 1 and st["rebuilds"] == 0
        g = same_as_fresh(f"move {k}")
    tiles = np.arange(0, 40, 3, dtype=np.int32)
    m = tile_mask(480, 270, tiles)
    _check(compare_frames(g, _oracle(hs, p, tiles=tiles), region=m))
    hs.switch(2, 3, False)
    ds.update(hs)
    assert ds.stats()["rebuilds"] == 1
    same_as_fresh("switch")
    hs.move_pos(2, 0, (0.1, 0.1, 0.1))
    ds.rebuild(hs)
    st = ds.stats()
    assert st["rebuilds"] == 2 and st["refits"] == 4
    same_as_fresh("rebuild")
