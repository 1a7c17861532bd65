#!/usr/bin/env python3
"""Benchmark: Mrays/s (primary + secondary) at 1080p on workload 2.

Workload 2: 1920x1080, 1 spp, depth 4; a 200,000-triangle noise-displaced
geodesic icosphere, 64 spheres and 4 point lights with shadow rays, seeded
and built through the Scene API; BVH built on the device once (scene stays
resident in HBM, as between frames of the same scene).

A step is one frame: every primary, shadow, reflection and refraction ray of
the reference recursion traced on the GPU, outputs written to device memory.
Ray counts are the reference-semantic counts accumulated by the kernels
(tested equal to the oracle's).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--engine auto|mega|wave]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def ray_bytes(rays, node_visits, tri_tests, sph_tests) -> float:
    """Algorithmic bytes of `rays` traced rays: 32 B ray read + 32 B hit/payload
    write per ray + 64 B per node visit + 48 B per triangle test + 16 B per
    sphere test, with the visit/test counts of the canonical oracle BVH."""
    return 64.0 * rays + 64.0 * node_visits + 48.0 * tri_tests + 16.0 * sph_tests


def fp64_counts(g: dict, o64: dict, W: int, H: int, tiles) -> list:
    """[pixels, id mismatches, sum of RGB absolute errors, max RGB error, pixels
    off by more than one step] of a device frame against the FP64 oracle on the
    rendered tiles."""
    import numpy as np
    m = np.zeros((H, W), bool)
    tx = (W + 63) // 64
    for t in tiles:
        bx, by = t % tx, t // tx
        m[by * 64:by * 64 + 64, bx * 64:bx * 64 + 64] = True
    ids_eq = (g["ids"] == o64["ids"]).all(-1)
    d = np.abs(g["rgb"].astype(np.int16) - o64["rgb"].astype(np.int16))
    return [int(m.sum()), int((~ids_eq & m).sum()), float(d[m].astype(np.float64).sum()), int(d[m].max()),
            int(((d.max(-1) > 1) & m).sum())]


def fp64_stats(c: list) -> dict:
    """fp64_counts (or their sums over frames) as the reported statistics:
    primary-id mismatch fraction, RGB mean / max absolute error (1/255 steps)
    and the fraction of pixels off by more than one step."""
    px = max(1, c[0])
    return {"pixels": c[0], "ids_mismatch": round(c[1] / px, 6), "rgb_mae": round(c[2] / (3 * px), 4),
            "rgb_max": c[3], "rgb_over_1": round(c[4] / px, 6)}


def fp64_error(g: dict, o64: dict, W: int, H: int, tiles) -> dict:
    """Pixel-error statistics of a device frame against the FP64 oracle on the
    rendered tiles."""
    return fp64_stats(fp64_counts(g, o64, W, H, tiles))


def sample_requests(hs, reqs, tiles_per: int = 4, seed: int = 7) -> list:
    """[(scene, width, height, tiles)]: the first request of each scene in
    `reqs` and up to `tiles_per` seeded 64x64 tiles of its frame."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for s in range(len(hs)):
        ks = [k for k, r in enumerate(reqs) if r[0] == s]
        if not ks:
            continue
        _, w, h = reqs[ks[0]][:3]
        tx, ty = (w + 63) // 64, (h + 63) // 64
        tiles = np.sort(rng.choice(tx * ty, size=min(tiles_per, tx * ty), replace=False)).astype(np.int32)
        out.append((s, w, h, tiles))
    return out


def _sample_what(sample, tiles_per: int = 4) -> str:
    return (f"the first request of each scene ({', '.join(f'{w}x{h}' for _, w, h, _ in sample)}), up to "
            f"{tiles_per} seeded 64x64 tiles each")


def request_errors(rt, hs, ds, reqs, got, ref) -> dict:
    """Pixel errors of the batch and serving workloads: every request's RGB in
    `got` (batched) against `ref` (the same request rendered alone), byte by
    byte, and the sampled requests (sample_requests), rendered alone with their
    ids, against the FP64 oracle on their sampled tiles."""
    import numpy as np
    from oracle import pyoracle as po
    same = sum(int(np.array_equal(a, b)) for a, b in zip(got, ref))
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    sample = sample_requests(hs, reqs)
    tot = [0, 0, 0.0, 0, 0]
    for s, w, h, tiles in sample:
        p = rt.params(w, h, max_level=hs[s].info.max_level)
        o64 = po.OracleScene(hs[s].desc_ptr).render(p, po.ORACLE_FP64, threads=threads, tiles=tiles)
        c = fp64_counts(ds[s].render(p), o64, w, h, tiles)
        tot = [tot[0] + c[0], tot[1] + c[1], tot[2] + c[2], max(tot[3], c[3]), tot[4] + c[4]]
    out = fp64_stats(tot)
    out["sample"] = _sample_what(sample) + ", rendered alone"
    return {"fp64_error": out, "requests_identical_to_alone": f"{same} of {len(got)}"}


def cpu_requests(rt, hs, reqs) -> dict:
    """The reference algorithm's port (the REF32 oracle) on the host cores over
    the sampled tiles of the sampled requests (sample_requests), one wall-clock
    run per request: Mrays/s."""
    from oracle import pyoracle as po
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    sample = sample_requests(hs, reqs)
    rays, secs, px = 0, 0.0, 0
    for s, w, h, tiles in sample:
        p = rt.params(w, h, max_level=hs[s].info.max_level)
        st = po.OracleScene(hs[s].desc_ptr).render(p, po.ORACLE_REF32, threads=threads, tiles=tiles,
                                                    want=("rgb",))["stats"]
        rays, secs, px = rays + st.rays(), secs + st.seconds, px + st.pixels
    return {"value": round(rays / max(secs, 1e-9) / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{_sample_what(sample)} ({px} pixels, {rays} reference rays, {secs:.1f} s); the rate "
                      "extrapolates to the whole workload"}


def traffic_model(tr) -> dict:
    """Per-ray algorithmic bytes by ray kind from an oracle traffic count (shadow
    rays whose outcome cannot change the colour are not traced and cost 0)."""
    ref_rays = tr.rays + tr.shadow_skipped  # reference-semantic rays of the sample
    shadow = ray_bytes(tr.shadow_rays, tr.shadow_node_visits, tr.shadow_tri_tests, tr.shadow_sph_tests)
    closest = ray_bytes(tr.rays - tr.shadow_rays, tr.node_visits - tr.shadow_node_visits,
                        tr.tri_tests - tr.shadow_tri_tests, tr.sph_tests - tr.shadow_sph_tests)
    return {
        "per_reference_ray": (shadow + closest) / max(1, ref_rays),
        "per_shadow_ray": shadow / max(1, tr.shadow_rays),
        "per_closest_ray": closest / max(1, tr.rays - tr.shadow_rays),
        "shadow_skipped_frac": tr.shadow_skipped / max(1, tr.shadow_rays + tr.shadow_skipped),
    }


# Launch tuning of the 1080p frame, chosen by rtg_autotune on MI355X: 16x4
# pixel blocks per wave, one 64-item chunk per grab, persistent grids at half
# the resident blocks (the render and side streams' kernels run side by side).
FRAME_TUNING = dict(tile_w=16, grab=1, grid_pct=50)


def pmc_traffic(engine: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, scaled by a
    plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_workload2_1080p.json")
    try:
        with open(path) as f:
            return json.load(f)[engine][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def pmc_summary(d: str, engine: str = "wave") -> dict:
    """HBM bytes per dispatch by kernel from the rocprofv3 --pmc CSVs under d:
    copy_fetch/ and copy_write/ (FETCH_SIZE, WRITE_SIZE of a --calibrate-copy
    run) and <engine>_fetch/, <engine>_write/ (the same counters over frames).
    Each counter is scaled by what it reads for the plain copy: 2^30 bytes read
    and written per copy dispatch."""
    import csv
    import glob
    import re
    from collections import defaultdict

    def per_kernel(sub, counter):
        total, disp = defaultdict(float), defaultdict(set)
        for path in glob.glob(os.path.join(d, sub, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r["Counter_Name"] != counter:
                        continue
                    m = re.search(r"(k_\w+|__amd_rocclr_\w+)", r["Kernel_Name"])
                    k = m.group(1) if m else r["Kernel_Name"]
                    total[k] += float(r["Counter_Value"])
                    disp[k].add(r["Dispatch_Id"])
        return {k: (total[k] / len(disp[k]), len(disp[k])) for k in total}

    copy_kib = (1 << 30) / 1024.0
    cf = per_kernel("copy_fetch", "FETCH_SIZE")
    cw = per_kernel("copy_write", "WRITE_SIZE")
    copy_k = [k for k in cf if "copy" in k.lower()]
    if not copy_k:
        raise RuntimeError(f"no copy kernel under {d}/copy_fetch")
    fetch_factor = copy_kib / cf[copy_k[0]][0]
    write_factor = copy_kib / cw[copy_k[0]][0]
    fetch, write = per_kernel(f"{engine}_fetch", "FETCH_SIZE"), per_kernel(f"{engine}_write", "WRITE_SIZE")
    out = {}
    for k in sorted(set(fetch) & set(write)):
        out[k] = {"hbm_bytes_per_launch": int(round((fetch[k][0] * fetch_factor + write[k][0] * write_factor) * 1024)),
                  "fetch_kib_raw": round(fetch[k][0], 1), "write_kib_raw": round(write[k][0], 1),
                  "dispatches": fetch[k][1]}
    return {engine: out, "calibration": {"fetch_factor": round(fetch_factor, 4), "write_factor": round(write_factor, 4),
                                         "copy_bytes": 1 << 30}}


def batch_scenes(rt):
    """The 8 resident scenes of the batch and serving workloads (triangle soups
    and height fields of 1k..1M triangles, 1-8 lights, depth 2-6)."""
    hs = [rt.HostScene.workload(3, index=i) for i in range(8)]
    return hs, [rt.DeviceScene(h) for h in hs]


def bench_batch(args) -> dict:
    """Workload 3: 256 requests (64x64 .. 1024x768) over the 8 resident scenes,
    rendered as one batch (one wavefront, replayed from a captured hipGraph),
    against the plain renderer (one frame per request)."""
    import raytrace_amd as rt
    rt.init(0)
    hs, ds = batch_scenes(rt)
    reqs = rt.gen_requests(256, 0.0, seed=3)
    items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level)) for s, w, h, _ in reqs]

    # The host result buffers are allocated once and reused by every batch, as
    # a server reuses its response buffers: fresh pages per batch would time
    # the host's page faults (76 MB of RGB per batch), not the renderer.
    outs, _ = rt.render_batch(items, ids=False)

    def timed_batches():
        for _ in range(max(2, args.warmup)):
            rt.render_batch(items, ids=False, out=outs)
        t0 = time.perf_counter()
        replays = 0
        dev_ms = 0.0
        for _ in range(args.steps):
            _, st = rt.render_batch(items, ids=False, out=outs)
            replays += st.graph_replays
            dev_ms += st.render_ms
        return (time.perf_counter() - t0) / args.steps, st, replays, dev_ms / args.steps

    default_tuning = rt.get_tuning()
    tune = None
    if not args.no_autotune:
        # the tuner's runs also warm the GPU, so both timings below start warm
        t0 = time.perf_counter()
        best, trials = rt.autotune(items, reps=3)
        tune_s = time.perf_counter() - t0
        rt.set_tuning(**default_tuning)
        default_s, _, _, _ = timed_batches()
        rt.set_tuning(**best)
        tune = {"best": best, "default": default_tuning, "seconds": round(tune_s, 2),
                "trials": [{**t, "device_ms": round(ms, 3)} for t, ms in trials],
                "default_ms_per_batch": round(default_s * 1e3, 3)}
    batch_s, st, replays, dev_ms = timed_batches()
    rays = st.rays_total
    pouts = [d.render(p, ids=False, tprim=False) for d, p in items]  # warm, and the reused host buffers
    t0 = time.perf_counter()
    prays = 0
    for (d, p), o in zip(items, pouts):
        prays += d.render(p, ids=False, tprim=False, out=o)["stats"].rays_total
    plain_s = time.perf_counter() - t0
    assert prays == rays
    errs = request_errors(rt, hs, ds, reqs, [o["rgb"] for o in outs], [o["rgb"] for o in pouts])
    cpu = None if args.no_cpu_baseline else cpu_requests(rt, hs, reqs)
    return {
        "metric": "Mrays/s (primary+secondary), 256 batched requests", "value": round(rays / batch_s / 1e6, 2),
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batch_s * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 requests 64x64..1024x768 over 8 resident scenes (1k..1M triangles, 1-8 lights, "
                               "depth 2-6), RGB to host", "host_buffers": "allocated once, reused by every batch and frame",
                   "rays_per_batch": rays,
                   "pixels_per_batch": sum(p.width * p.height for _, p in items),
                   "scene_triangles": [int(h.info.n_tris) for h in hs], "graph_replayed_steps": replays,
                   "device_ms_per_batch": round(dev_ms, 3),
                   "tuning": rt.get_tuning()},
        "autotune": tune,
        "plain": {"ms_per_batch": round(plain_s * 1e3, 3), "mrays_per_s": round(rays / plain_s / 1e6, 2),
                  "what": "the same 256 requests, one rtg_render each"},
        "speedup_vs_plain": round(plain_s / batch_s, 2),
        "cpu_baseline": cpu,
        **errs,
    }


def bench_serve(args) -> dict:
    """Workload 4: Poisson arrivals at each rate for `--duration` seconds over the
    8 resident scenes; the continuous-batching loop against the plain renderer
    (one frame per request, first come first served) on the same arrivals."""
    import raytrace_amd as rt
    rt.init(0)
    hs, ds = batch_scenes(rt)
    rates = [float(r) for r in args.rates.split(",")]
    # warm-up, untimed: a full run of each mode at the first rate (slot
    # streams, every slot's workspace grown to the request sizes, kernel code
    # loading), as a server is warm before it takes traffic
    warm = rt.gen_requests(max(8, int(rates[0] * args.duration)), rates[0], seed=40)
    warm_items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level), t) for s, w, h, t in warm]
    for mode in (rt._capi.RTG_SERVE_BATCHED, rt._capi.RTG_SERVE_PLAIN):
        rt.serve(warm_items, mode=mode, slots=args.slots)
    rows = []
    top_rgb = {}  # each mode's results at the top rate, for the pixel errors
    for rate in rates:
        n = max(8, int(rate * args.duration))
        reqs = rt.gen_requests(n, rate, seed=4)
        items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level), t) for s, w, h, t in reqs]
        row = {"rate": rate, "requests": n}
        for name, mode in (("batched", rt._capi.RTG_SERVE_BATCHED), ("plain", rt._capi.RTG_SERVE_PLAIN)):
            res, _, st = rt.serve(items, mode=mode, slots=args.slots)
            top_rgb[name] = (reqs, res) if rate == rates[-1] else None
            row[name] = {"p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
                         "req_per_s": round(st.req_per_s, 1), "mrays_per_s": round(st.mrays_per_s, 1),
                         "batches": st.batches, "max_in_flight": st.max_in_flight}
            sys.stderr.write(f"rate {rate:g}/s {name}: {row[name]}\n")
            sys.stderr.flush()
        rows.append(row)
    # Capacity: the batched loop keeps up with every rate of the sweep, so its
    # throughput there is the offered rate; higher offered rates (batched
    # only: the plain renderer saturates within the sweep) find its ceiling.
    cap_rows = []
    for rate in [float(r) for r in args.capacity_rates.split(",") if r]:
        n = max(8, int(rate * args.duration))
        reqs = rt.gen_requests(n, rate, seed=4)
        items = [(ds[s], rt.params(w, h, max_level=hs[s].info.max_level), t) for s, w, h, t in reqs]
        _, _, st = rt.serve(items, mode=rt._capi.RTG_SERVE_BATCHED, slots=args.slots)
        cap_rows.append({"rate": rate, "requests": n, "batched": {
            "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "req_per_s": round(st.req_per_s, 1),
            "mrays_per_s": round(st.mrays_per_s, 1), "batches": st.batches, "max_in_flight": st.max_in_flight}})
        sys.stderr.write(f"rate {rate:g}/s batched: {cap_rows[-1]['batched']}\n")
        sys.stderr.flush()
    top = rows[-1]
    # the plain renderer's highest throughput, and the batched loop's highest
    # throughput at a p99 no worse than the plain renderer's there
    best_p = max((r["plain"] for r in rows), key=lambda b: b["req_per_s"])
    ok_b = [r["batched"] for r in rows + cap_rows if r["batched"]["p99_ms"] <= best_p["p99_ms"]]
    best_b = max(ok_b or [top["batched"]], key=lambda b: b["req_per_s"])
    # the batched loop's results against the plain renderer's (each request
    # alone) at the top rate, byte by byte, and sampled requests against FP64
    errs = request_errors(rt, hs, ds, top_rgb["batched"][0], top_rgb["batched"][1], top_rgb["plain"][1])
    cpu = None
    if not args.no_cpu_baseline:
        # the port's ray rate as requests/s of this workload (its rays per request)
        c = cpu_requests(rt, hs, top_rgb["batched"][0])
        per_req = top["batched"]["mrays_per_s"] * 1e6 / max(1e-9, top["batched"]["req_per_s"])
        cpu = {**c, "value": round(c["value"] * 1e6 / per_req, 3), "unit": "requests/s",
               "mrays_per_s": c["value"], "rays_per_request": round(per_req)}
    return {
        "metric": "serve-loop p50/p99 frame latency and throughput", "value": top["batched"]["req_per_s"],
        "unit": "requests/s at the highest offered rate of the sweep", "n_gpus": 1, "steps": len(rates), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Poisson arrivals at {args.rates} req/s, {args.duration:g} s each, requests 64x64.."
                               "1024x768 over 8 resident scenes, RGB to host", "slots": args.slots,
                   "capacity_rates": args.capacity_rates,
                   "host_buffers": "result arrays written before each run starts (pooled response buffers)",
                   "warmup": "one untimed run of each mode at the first rate (other arrivals) before the sweep"},
        "rates": rows,
        "capacity": cap_rows,
        "throughput_vs_plain_at_top_rate": round(top["batched"]["req_per_s"] / max(1e-9, top["plain"]["req_per_s"]), 2),
        "max_throughput": {"plain": best_p, "batched_at_no_worse_p99": best_b,
                           "ratio": round(best_b["req_per_s"] / max(1e-9, best_p["req_per_s"]), 2)},
        "cpu_baseline": cpu,
        **errs,
    }


def bench_stress(args) -> dict:
    """Workload 5: 3840x2160, 4 jittered samples, depth 8, refraction; a
    4,188,162-triangle terrain + 256 sphere meshes (2,048 triangles each),
    16 lights.  Frames into device buffers timed by wall clock; pixel errors
    against the FP64 oracle and parity against the reference algorithm on
    tile samples."""
    import numpy as np
    import torch  # before librtg initialises HIP: torch then finds the device

    import raytrace_amd as rt
    from oracle import pyoracle as po
    from raytrace_amd import _capi
    rt.init(0)
    log = lambda m: (sys.stderr.write(m + "\n"), sys.stderr.flush())
    hs = rt.HostScene.workload(5)
    W, H = hs.info.width, hs.info.height
    log(f"workload 5: {hs.info.n_tris} triangles, {W}x{H}, building")
    ds = rt.DeviceScene(hs)
    p = rt.params(W, H, max_level=8, spp=4, seed=5)
    # frames into device buffers on a caller stream, as the 1080p line: the
    # wall time of the timed frames is the rate, outputs stay in HBM
    lib = _capi.load()
    dev = torch.device("cuda:0")
    rgb_d = torch.empty((H, W, 3), dtype=torch.uint8, device=dev)
    counts = torch.zeros(4, dtype=torch.int64, device=dev)
    stream = torch.cuda.Stream(device=dev)
    sp = C.c_void_p(stream.cuda_stream)

    def frame(pp):
        rc = lib.rtg_render_device(ds.h, C.byref(pp), C.c_void_p(rgb_d.data_ptr()), None, None,
                                   C.c_void_p(counts.data_ptr()), sp)
        if rc != 0:
            raise RuntimeError(f"rtg_render_device failed: {_capi.last_error()}")

    n_warm, n_steps = max(1, args.warmup), max(1, args.steps)
    for k in range(n_warm):
        frame(p)
        stream.synchronize()
        log(f"warmup frame {k}")
    t0 = time.perf_counter()
    for k in range(n_steps):
        frame(p)
        stream.synchronize()
        log(f"frame {k}: {time.perf_counter() - t0:.2f} s")
    wall_s = (time.perf_counter() - t0) / n_steps
    per_type = [int(x) for x in counts.tolist()]
    rays = sum(per_type)
    # one more frame with kernel timing: the shadow rays k_occlude traced
    rt.kernel_times()
    frame(rt.params(W, H, max_level=8, spp=4, seed=5, flags=rt.RTG_FLAG_KERNEL_TIMES))
    stream.synchronize()
    kt = rt.kernel_times()
    traced_shadow = kt["k_occlude"][1] if "k_occlude" in kt else per_type[1]
    out = ds.render(p)  # host copy with ids, for the oracle checks
    st = out["stats"]
    assert st.rays_total == rays
    osc = po.OracleScene(hs.desc_ptr)
    tx, ty = (W + 63) // 64, (H + 63) // 64
    rng = np.random.default_rng(5)
    tiles = np.sort(rng.choice(tx * ty, size=(tx * ty) // 64, replace=False)).astype(np.int32)
    # host cores this process may use, at most the 16 of a one-GPU share
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    log(f"FP64 oracle on {len(tiles)} tiles")
    o64 = osc.render(p, po.ORACLE_FP64, threads=threads, tiles=tiles)
    fp64 = fp64_error(out, o64, W, H, tiles)
    ref_tiles = tiles[:: max(1, len(tiles) // 3)][:3]
    log(f"reference-algorithm oracle on {len(ref_tiles)} tiles")
    o32 = osc.render(p, po.ORACLE_REF32, threads=threads, tiles=ref_tiles)
    mr = np.zeros((H, W), bool)
    for t in ref_tiles:
        bx, by = t % tx, t // tx
        mr[by * 64:by * 64 + 64, bx * 64:bx * 64 + 64] = True
    eq = (out["ids"] == o32["ids"]).all(-1)
    band = (o32["flags"] & 3) != 0
    dr = np.abs(out["rgb"].astype(np.int16) - o32["rgb"].astype(np.int16)).max(-1)
    parity = {"pixels": int(mr.sum()), "ids_match": round(float(eq[mr].mean()), 6),
              "ids_mismatch_outside_band": int((~eq & ~band & mr).sum()), "rgb_over_1": int(((dr > 1) & eq & mr).sum())}
    # the reference algorithm's port on the host cores: that same render, timed
    s32 = o32["stats"]
    cpu = {"value": round(s32.rays() / s32.seconds / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"the {len(ref_tiles)} 64x64 tiles of the parity check ({s32.pixels} pixels, {s32.rays()} "
                     f"reference rays, {s32.seconds:.1f} s), one wall-clock run; the rate extrapolates to the frame"}
    rays_traced = rays - per_type[1] + traced_shadow
    return {
        "metric": "Mrays/s (primary+secondary), 4K stress frame", "value": round(rays / wall_s / 1e6, 2),
        "unit": "Mrays/s", "n_gpus": 1, "steps": n_steps, "warmup": n_warm,
        "ms_per_step": round(wall_s * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3840x2160, 4 spp jittered, depth 8, refraction IOR 1.3-1.7, 4,188,162-tri terrain + "
                               "256 x 2,048-tri sphere meshes, 16 lights",
                   "timing": "wall clock of frames into device buffers (rtg_render_device), outputs in HBM",
                   "rays_per_frame": rays, "rays_traced_per_frame": rays_traced,
                   "traced_mrays_per_s": round(rays_traced / wall_s / 1e6, 2),
                   "shadow_rays_skipped_frac": round(1.0 - traced_shadow / max(1, per_type[1]), 4),
                   "rays_by_type": {"primary": st.rays_primary, "shadow": st.rays_shadow, "reflect": st.rays_reflect,
                                    "refract": st.rays_refract},
                   "bvh_build_ms": round(ds.build_ms, 2), "scene_bytes": ds.nbytes, "triangles": int(hs.info.n_tris)},
        "cpu_baseline": cpu, "fp64_error": fp64, "parity_vs_reference_algorithm": parity,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--engine", default="auto", choices=["mega", "wave", "auto"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-tiles", type=int, default=16, help="1/N of the 64x64 tiles for the CPU baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline: median of this many runs")
    ap.add_argument("--workload", type=int, default=2, choices=[2, 3, 4, 5],
                    help="2: the 1080p frame (default line); 3: 256 batched requests; 4: serving sweep; "
                         "5: 4K stress frame")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="workload 4: offered rates (requests/s)")
    ap.add_argument("--duration", type=float, default=2.0, help="workload 4: seconds of arrivals per rate")
    ap.add_argument("--slots", type=int, default=2, help="workload 4: concurrent batches")
    ap.add_argument("--capacity-rates", default="4000,8000",
                    help="workload 4: offered rates beyond the sweep, batched loop only (its throughput ceiling)")
    ap.add_argument("--no-autotune", action="store_true", help="workload 3: keep the default launch tuning")
    ap.add_argument("--autotune", action="store_true", help="workload 2: autotune the launch tuning first")
    ap.add_argument("--default-tuning", action="store_true",
                    help="workload 2: the library's default launch tuning instead of the frame tuning")
    ap.add_argument("--one-stream", action="store_true",
                    help="workload 2: every kernel on the render stream (RTG_FLAG_ONE_STREAM), for counter collection")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference for FETCH_SIZE / WRITE_SIZE scaling)")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="only print HBM bytes per dispatch by kernel from the rocprofv3 --pmc runs under DIR "
                         "(copy_fetch, copy_write, wave_fetch, wave_write)")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary), indent=1))
        return 0
    if args.calibrate_copy:
        import torch
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes": 1 << 30, "calls": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(frames and requests shard across GPUs without exchange; see the design notes)", file=sys.stderr)
        return 2
    if args.workload == 3:
        print(json.dumps(bench_batch(args)))
        return 0
    if args.workload == 4:
        print(json.dumps(bench_serve(args)))
        return 0
    if args.workload == 5:
        print(json.dumps(bench_stress(args)))
        return 0

    import numpy as np
    import torch

    import raytrace_amd as rt
    from raytrace_amd import _capi

    engine = {"mega": rt.RTG_ENGINE_MEGAKERNEL, "wave": rt.RTG_ENGINE_WAVEFRONT, "auto": rt.RTG_ENGINE_AUTO}[args.engine]
    eng_name = "mega" if args.engine == "mega" else "wave"  # AUTO selects the wavefront engine
    rt.init(0)
    lib = _capi.load()
    W, H, DEPTH = 1920, 1080, 4
    hs = rt.HostScene.workload(2)
    ds = rt.DeviceScene(hs)
    build_ms = ds.build_ms
    flags = _capi.RTG_FLAG_ONE_STREAM if args.one_stream else 0
    p = rt.params(W, H, max_level=DEPTH, engine=engine, flags=flags)

    dev = torch.device("cuda:0")
    rgb = torch.empty((H, W, 3), dtype=torch.uint8, device=dev)
    counts = torch.zeros(4, dtype=torch.int64, device=dev)
    stream = torch.cuda.Stream(device=dev)
    sp = C.c_void_p(stream.cuda_stream)

    def step():
        rc = lib.rtg_render_device(ds.h, C.byref(p), C.c_void_p(rgb.data_ptr()), None, None,
                                   C.c_void_p(counts.data_ptr()), sp)
        if rc != 0:
            raise RuntimeError(f"rtg_render_device failed: {_capi.last_error()}")

    # Launch tuning of the frame path, chosen by rtg_autotune on this frame on
    # MI355X (persistent grids at 75% of the resident blocks, so the render
    # and side streams' kernels run side by side), installed as a caller that
    # renders this frame repeatedly would; --autotune re-runs the tuner,
    # --default-tuning keeps the library default.
    tune = None
    if args.autotune:
        default_tuning = rt.get_tuning()
        best, trials = rt.autotune([(ds, p)], reps=3)
        tune = {"best": best, "default": default_tuning,
                "trials": [{**t, "device_ms": round(ms, 3)} for t, ms in trials]}
    elif not args.default_tuning and eng_name == "wave":
        rt.set_tuning(**FRAME_TUNING)
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    rays_per_frame = int(counts.sum().item())
    per_type = [int(x) for x in counts.tolist()]

    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1000.0 / args.steps
    value = rays_per_frame / (ms_per_step / 1000.0) / 1e6

    # Per-kernel device time: the same K frames again, with HIP events around
    # every launch on the render stream (RTG_FLAG_KERNEL_TIMES).
    pk = rt.params(W, H, max_level=DEPTH, engine=engine, flags=rt.RTG_FLAG_KERNEL_TIMES | flags)
    rt.kernel_times()
    for _ in range(args.steps):
        rc = lib.rtg_render_device(ds.h, C.byref(pk), C.c_void_p(rgb.data_ptr()), None, None,
                                   C.c_void_p(counts.data_ptr()), sp)
        if rc != 0:
            raise RuntimeError(f"rtg_render_device failed: {_capi.last_error()}")
    stream.synchronize()
    kt = rt.kernel_times()
    dom = max(kt, key=lambda k: kt[k][2])
    launches, items, ms = kt[dom]
    # The same again with every kernel on the render stream: the dominant
    # kernel's launch time when it has the GPU to itself.
    alone_ms = None
    if eng_name == "wave" and not args.one_stream:
        pa = rt.params(W, H, max_level=DEPTH, engine=engine,
                       flags=rt.RTG_FLAG_KERNEL_TIMES | _capi.RTG_FLAG_ONE_STREAM)
        rt.kernel_times()
        for _ in range(args.steps):
            rc = lib.rtg_render_device(ds.h, C.byref(pa), C.c_void_p(rgb.data_ptr()), None, None,
                                       C.c_void_p(counts.data_ptr()), sp)
            if rc != 0:
                raise RuntimeError(f"rtg_render_device failed: {_capi.last_error()}")
        stream.synchronize()
        ka = rt.kernel_times()
        if dom in ka:
            alone_ms = ka[dom][2] / ka[dom][0]

    # algorithmic traffic of the frame's ray set, canonical oracle BVH, tile sample
    from oracle import pyoracle as po
    osc = po.OracleScene(hs.desc_ptr)
    tx, ty = (W + 63) // 64, (H + 63) // 64
    rng = np.random.default_rng(11)
    # host cores this process may use, at most the 16 of a one-GPU share
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    traffic_tiles = np.sort(rng.choice(tx * ty, size=(tx * ty) // 16, replace=False)).astype(np.int32)
    tm = traffic_model(osc.traffic(p, threads=threads, tiles=traffic_tiles))
    # CPU baseline sample: a seeded 1/N of the frame's tiles (N = 16: about
    # 1.3 s per run of the reference-algorithm port on 16 host threads)
    tiles = np.sort(np.random.default_rng(12).choice(tx * ty, size=max(1, (tx * ty) // args.cpu_tiles),
                                                     replace=False)).astype(np.int32)
    def kernel_roofline(name):
        """Algorithmic GB/s of one kernel's average launch: its work units per
        launch x the canonical-BVH bytes per unit, over its average duration."""
        n_l, n_items, k_ms = kt[name]
        if name == "k_occlude":
            u, b = "traced shadow ray", tm["per_shadow_ray"]
        elif name == "k_extend":
            u, b = "closest-hit ray", tm["per_closest_ray"]
        else:  # the megakernel traces the whole frame
            u, b = "reference ray", tm["per_reference_ray"]
            n_items = rays_per_frame * n_l
        per_launch = n_items * b / n_l
        avg = k_ms / n_l
        ach = per_launch / (avg / 1000.0) / 1e9
        return {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4), "avg_launch_ms": round(avg, 4),
                "launches_per_frame": n_l // args.steps, "work_unit": u, "units_per_launch": round(n_items / n_l),
                "bytes_per_unit": round(b, 1)}, per_launch, avg, u, b, n_items

    _, bytes_per_launch, avg_ms, unit, per_unit, items = kernel_roofline(dom)
    achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9
    # rays the device traced: the reference-semantic count less the shadow rays
    # whose outcome cannot change the colour (not traced, still counted)
    traced_shadow = kt["k_occlude"][1] // args.steps if "k_occlude" in kt else per_type[1]
    rays_traced = rays_per_frame - per_type[1] + traced_shadow
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(eng_name, dom),
                "kernel": dom, "launches_per_frame": launches // args.steps, "avg_launch_ms": round(avg_ms, 4),
                "work_unit": unit, "units_per_launch": round(items / launches), "bytes_per_unit": round(per_unit, 1),
                "alone": None if alone_ms is None else {
                    "avg_launch_ms": round(alone_ms, 4),
                    "frac": round(bytes_per_launch / (alone_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                    "what": "the same frames with every kernel on the render stream (RTG_FLAG_ONE_STREAM)"},
                "frame": {"device_ms": round(frame_ms, 4),
                          "achieved_GBs": round(rays_per_frame * tm["per_reference_ray"] / frame_ms / 1e6, 1),
                          "bytes_per_reference_ray": round(tm["per_reference_ray"], 1),
                          "shadow_rays_skipped_frac": round(tm["shadow_skipped_frac"], 4)},
                "kernels_ms_per_frame": {k: round(v[2] / args.steps, 4) for k, v in sorted(kt.items())},
                "kernels": {k: kernel_roofline(k)[0] for k in ("k_extend", "k_occlude") if k in kt}}

    # pixel errors of the frame against the FP64 oracle, on the traffic tile sample
    g = ds.render(rt.params(W, H, max_level=DEPTH, engine=engine))
    o64 = osc.render(p, po.ORACLE_FP64, threads=threads, tiles=traffic_tiles)
    fp64 = fp64_error(g, o64, W, H, traffic_tiles)
    fp64["sample"] = f"seeded {len(traffic_tiles)} of {tx * ty} 64x64 tiles"

    cpu = None
    if not args.no_cpu_baseline:
        def timed(mode, nthreads, runs):
            """Median Mrays/s over `runs` wall-clock renders of the tile sample."""
            rates = []
            for _ in range(runs):
                st = osc.render(p, mode, threads=nthreads, tiles=tiles, want=("rgb",))["stats"]
                rates.append(st.rays() / st.seconds / 1e6)
            return float(np.median(rates)), [round(r, 4) for r in rates], st

        runs = args.cpu_runs
        v_all, r_all, st = timed(po.ORACLE_REF32, threads, runs)
        v_8, r_8, _ = timed(po.ORACLE_REF32, 8, runs)
        v_o, r_o, _ = timed(po.ORACLE_FP64, threads, runs)
        cpu = {"value": round(v_all, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"seeded {len(tiles)} of {tx * ty} 64x64 tiles of the 1080p frame ({st.pixels} pixels, "
                         f"{st.rays()} reference rays per run), median of {runs} wall-clock runs; the rate "
                         "extrapolates to the full frame",
               "runs": r_all,
               "tnum_8": {"value": round(v_8, 3), "cores": 8, "runs": r_8,
                          "what": "the same port at the reference's default of 8 worker threads"},
               "oracle": {"value": round(v_o, 3), "cores": threads, "runs": r_o,
                          "what": "the FP64 BVH oracle (not the reference algorithm), same sample"}}

    line = {
        "metric": "Mrays/s (primary+secondary) at 1080p", "value": round(value, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1920x1080 1spp depth 4, 200k-tri displaced icosphere + 64 spheres, 4 point lights",
                   "engine": {"mega": "megakernel", "wave": "wavefront"}[eng_name], "rays_per_frame": rays_per_frame,
                   "rays_by_type": dict(zip(["primary", "shadow", "reflect", "refract"], per_type)),
                   "rays_traced_per_frame": rays_traced,
                   "traced_mrays_per_s": round(rays_traced / (ms_per_step / 1000.0) / 1e6, 2),
                   "shadow_rays_skipped_frac": round(1.0 - traced_shadow / max(1, per_type[1]), 4),
                   "bvh_build_ms": round(build_ms, 3), "tuning": rt.get_tuning(),
                   "side_stream": not args.one_stream},
        "roofline": roofline, "cpu_baseline": cpu, "fp64_error": fp64,
    }
    if tune is not None:
        line["autotune"] = tune
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
