"""raytrace_amd — MI355X-native render path of the RayTrace Whitted renderer.

Python face of librtg (include/rtg.h, include/rtg_host.h):

    import raytrace_amd as rt
    rt.init(0)
    hs = rt.HostScene.workload(1)             # seeded scene, built through the Scene API
    ds = rt.DeviceScene(hs)                   # resident in HBM, BVH built on the device
    out = ds.render(rt.params(640, 480, max_level=3))
    out["rgb"]  # (H, W, 3) uint8, row 0 = bottom image row as in the reference

There is no CPU fallback: every render runs the HIP kernels of librtg.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import (RTG_ENGINE_AUTO, RTG_ENGINE_MEGAKERNEL, RTG_ENGINE_WAVEFRONT, RTG_FLAG_EAGER,
                    RTG_FLAG_FP64_RETEST, RTG_FLAG_KERNEL_TIMES, RTG_FLAG_ONE_STREAM,
                    RTG_MODE_RAYTRACE, RtgError, check, rtg_gen_info, rtg_gen_opts, rtg_params, rtg_stats)

__all__ = [
    "init", "shutdown", "params", "HostScene", "DeviceScene", "RtgError", "RTG_ENGINE_AUTO",
    "RTG_ENGINE_MEGAKERNEL", "RTG_ENGINE_WAVEFRONT", "RTG_FLAG_EAGER", "RTG_FLAG_FP64_RETEST",
    "RTG_FLAG_KERNEL_TIMES", "RTG_FLAG_ONE_STREAM", "RTG_MODE_RAYTRACE",
]


def init(device: int = 0) -> None:
    check(_capi.load().rtg_init(device), "rtg_init")


def shutdown() -> None:
    check(_capi.load().rtg_shutdown(), "rtg_shutdown")


def params(width: int, height: int, max_level: int = 1, spp: int = 1, tile_compat: int = 0,
           engine: int = RTG_ENGINE_AUTO, flags: int = 0, seed: int = 0, mode: int = RTG_MODE_RAYTRACE) -> rtg_params:
    return rtg_params(mode=mode, max_level=max_level, width=width, height=height, spp=spp, tile_compat=tile_compat,
                      engine=engine, flags=flags, seed=seed)


class HostScene:
    """A reference-API Scene (raytrace_amd/host) owned through the C ABI."""

    def __init__(self, handle: C.c_void_p, info: rtg_gen_info | None = None):
        self._lib = _capi.load()
        self.h = handle
        self.info = info

    @classmethod
    def new(cls) -> "HostScene":
        lib = _capi.load()
        h = C.c_void_p()
        check(lib.rtg_host_scene_new(C.byref(h)), "rtg_host_scene_new")
        return cls(h)

    @classmethod
    def workload(cls, config: int, seed: int = 0, width: int = 0, height: int = 0, detail: int = 0,
                 n_spheres: int = -1, n_lights: int = -1, index: int = 0, max_level: int = -1) -> "HostScene":
        lib = _capi.load()
        o = rtg_gen_opts(seed=seed, width=width, height=height, detail=detail, n_spheres=n_spheres,
                         n_lights=n_lights, index=index, max_level=max_level)
        h = C.c_void_p()
        info = rtg_gen_info()
        check(lib.rtg_host_scene_config(config, C.byref(o), C.byref(h), C.byref(info)), "rtg_host_scene_config")
        return cls(h, info)

    @property
    def desc_ptr(self) -> int:
        p = self._lib.rtg_host_scene_desc(self.h)
        if not p:
            raise RtgError(-1, "rtg_host_scene_desc")
        return p

    def close(self) -> None:
        if self.h:
            self._lib.rtg_host_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # Scene API (Scene.h:33-45)
    def add_sphere(self, r: float) -> int:
        return check(self._lib.rtg_host_add_sphere(self.h, r), "AddSphere")

    def add_plane(self) -> int:
        return check(self._lib.rtg_host_add_plane(self.h), "AddPlane")

    def add_cube(self, length: float) -> int:
        return check(self._lib.rtg_host_add_cube(self.h, length), "AddCube")

    def add_ballplane(self, r: float) -> int:
        return check(self._lib.rtg_host_add_ballplane(self.h, r), "AddBallPlane")

    def add_model_obj(self, obj_path: str, mtl_path: str) -> int:
        return check(self._lib.rtg_host_add_model_obj(self.h, obj_path.encode(), mtl_path.encode()), "AddModel")

    def add_light(self, type_: int, comp, atte=(1.0, 0.0, 0.0, 1.0)) -> int:
        return check(self._lib.rtg_host_add_light(self.h, type_, _capi.floats(*comp), _capi.floats(*atte)),
                     "AddLight")

    def move_pos(self, type_: int, num: int, v) -> None:
        check(self._lib.rtg_host_move_pos(self.h, type_, num, _capi.floats(*v)), "MovePos")

    def switch(self, type_: int, num: int, show: bool) -> int:
        return check(self._lib.rtg_host_switch(self.h, type_, num, int(show)), "Switch")

    def set_material(self, num: int, ka, kd, ks, shiness: float, reflect: float = 0.0, refract: float = 0.0,
                     rfr: float = 1.0) -> None:
        """Scene::ChgMtl(num, Material) (Scene.cpp:248-280) with the given components."""
        m = _capi.rtg_material()
        m.ambient, m.diffuse, m.specular = (_capi.rtg_vec4(*ka, 1.0), _capi.rtg_vec4(*kd, 1.0),
                                            _capi.rtg_vec4(*ks, 1.0))
        m.emission = _capi.rtg_vec4(0, 0, 0, 1)
        m.shiness, m.reflect, m.refract, m.rfr, m.tex = shiness, reflect, refract, rfr, -1
        check(self._lib.rtg_host_set_material(self.h, num, C.byref(m)), "ChgMtl")

    def set_position(self, num: int, pos) -> None:
        check(self._lib.rtg_host_set_position(self.h, num, _capi.floats(*pos)), "position")

    def set_light_position(self, num: int, pos) -> None:
        check(self._lib.rtg_host_set_light_position(self.h, num, _capi.floats(*pos)), "Light::SetProperty")

    def set_env_light(self, env) -> None:
        check(self._lib.rtg_host_set_env_light(self.h, _capi.floats(*env)), "EnvLight")

    def resize(self, w: int, h: int) -> None:
        check(self._lib.rtg_host_resize(self.h, w, h), "Camera::resize")

    def raytrace(self, type_: int = RTG_MODE_RAYTRACE, max_level: int = 1, frames: int = 1, with_uploads=False):
        """RayTracer::start + wait for isFinish, `frames` times on one RayTracer;
        returns (2048x2048x3 buffer of the last frame, its useTime[, scene uploads])."""
        out = np.full((2048 * 2048 * 3,), 127, np.uint8)
        t = C.c_double()
        up = C.c_uint64()
        check(self._lib.rtg_host_raytrace(self.h, type_, max_level, frames, out.ctypes.data_as(C.c_void_p),
                                          C.byref(t), C.byref(up)), "RayTracer::start")
        return (out, t.value, up.value) if with_uploads else (out, t.value)

    def raytrace_stopped(self, type_: int, max_level: int, stop_after_s: float):
        """RayTracer::start, RayTracer::stop after `stop_after_s` seconds, then wait
        for isFinish; returns (2048x2048x3 buffer, status, finished before the stop)."""
        out = np.full((2048 * 2048 * 3,), 127, np.uint8)
        fin = C.c_int()
        rc = self._lib.rtg_host_raytrace_stop(self.h, type_, max_level, float(stop_after_s),
                                              out.ctypes.data_as(C.c_void_p), C.byref(fin))
        return out, rc, bool(fin.value)


class DeviceScene:
    """A scene resident in HBM (rtg_scene_create)."""

    def __init__(self, host: HostScene):
        self._lib = _capi.load()
        self.h = C.c_void_p()
        check(self._lib.rtg_scene_create(C.c_void_p(host.desc_ptr), C.byref(self.h)), "rtg_scene_create")

    def update(self, host: HostScene) -> None:
        check(self._lib.rtg_scene_update(self.h, C.c_void_p(host.desc_ptr)), "rtg_scene_update")

    def rebuild(self, host: HostScene) -> None:
        """rtg_scene_rebuild: update with a fresh BVH (no refit)."""
        check(self._lib.rtg_scene_rebuild(self.h, C.c_void_p(host.desc_ptr)), "rtg_scene_rebuild")

    def stats(self) -> dict:
        """rtg_scene_stats: size and what the updates did."""
        info = _capi.rtg_scene_info()
        check(self._lib.rtg_scene_stats(self.h, C.byref(info)), "rtg_scene_stats")
        return {k: getattr(info, k) for k, _ in info._fields_}

    @property
    def build_ms(self) -> float:
        v = C.c_double()
        check(self._lib.rtg_scene_build_ms(self.h, C.byref(v)), "rtg_scene_build_ms")
        return v.value

    @property
    def nbytes(self) -> int:
        v = C.c_uint64()
        check(self._lib.rtg_scene_bytes(self.h, C.byref(v)), "rtg_scene_bytes")
        return v.value

    def render(self, p: rtg_params, ids: bool = True, tprim: bool = True, out: dict | None = None) -> dict:
        """rtg_render into new arrays, or into the arrays of `out` (a previous
        result of the same shape, reused: no fresh host pages per frame)."""
        W, H = p.width, p.height
        if out is not None:
            rgb, idb, tb = out["rgb"], out["ids"] if ids else None, out["t"] if tprim else None
            _check_out(rgb, (H, W, 3), np.uint8, "rgb")
            _check_out(idb, (H, W, 2), np.int32, "ids", ids)
            _check_out(tb, (H, W), np.float32, "t", tprim)
        else:
            rgb = np.empty((H, W, 3), np.uint8)
            idb = np.empty((H, W, 2), np.int32) if ids else None
            tb = np.empty((H, W), np.float32) if tprim else None
        st = rtg_stats()
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        check(self._lib.rtg_render(self.h, C.byref(p), ptr(rgb), ptr(idb), ptr(tb), C.byref(st)), "rtg_render")
        return {"rgb": rgb, "ids": idb, "t": tb, "stats": st}

    def close(self) -> None:
        if self.h:
            self._lib.rtg_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_times() -> dict:
    """Per-kernel device time of the renders issued with RTG_FLAG_KERNEL_TIMES
    since the previous call: {name: (launches, items, ms)}."""
    lib = _capi.load()
    buf = (_capi.rtg_kernel_time * 16)()
    n = C.c_int()
    check(lib.rtg_kernel_times(buf, 16, C.byref(n)), "rtg_kernel_times")
    return {buf[i].name.decode(): (buf[i].launches, buf[i].items, buf[i].ms) for i in range(min(n.value, 16))}



def _check_out(a, shape, dtype, what: str, wanted: bool = True) -> None:
    if not wanted:
        return
    if a is None or a.shape != shape or a.dtype != dtype or not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError(f"out[{what!r}] must be a writeable C-contiguous {np.dtype(dtype).name} array of shape {shape}")


def render_batch(items, ids: bool = True, eager: bool = False, out: list | None = None):
    """rtg_render_batch over [(DeviceScene, rtg_params), ...]: one wavefront for
    all requests.  Returns ([{"rgb", "ids"} per request], rtg_stats).  `out`: the
    result list of a previous batch of the same shapes, whose arrays are
    reused (no fresh host pages per batch)."""
    lib = _capi.load()
    n = len(items)
    if out is not None and len(out) != n:
        raise ValueError(f"out holds {len(out)} results for {n} requests")
    reqs = (_capi.rtg_request * max(1, n))()
    outs = []
    for k, (ds, p) in enumerate(items):
        q = rtg_params()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(rtg_params))
        if eager:
            q.flags |= _capi.RTG_FLAG_EAGER
        if out is not None:
            rgb, idb = out[k]["rgb"], out[k]["ids"] if ids else None
            _check_out(rgb, (p.height, p.width, 3), np.uint8, "rgb")
            _check_out(idb, (p.height, p.width, 2), np.int32, "ids", ids)
        else:
            rgb = np.empty((p.height, p.width, 3), np.uint8)
            idb = np.empty((p.height, p.width, 2), np.int32) if ids else None
        reqs[k].scene = ds.h
        reqs[k].params = q
        reqs[k].rgb = rgb.ctypes.data_as(C.POINTER(C.c_uint8))
        reqs[k].ids = idb.ctypes.data_as(C.POINTER(C.c_int32)) if ids else None
        outs.append({"rgb": rgb, "ids": idb})
    st = rtg_stats()
    check(lib.rtg_render_batch(reqs, n, C.byref(st)), "rtg_render_batch")
    return outs, st


def _tuning_dict(t) -> dict:
    return {"tile_w": t.tile_w, "grab": t.grab, "grid_pct": t.grid_pct}


def get_tuning() -> dict:
    """Launch tuning of the wavefront engine in use: {tile_w, grab, grid_pct}."""
    lib = _capi.load()
    t = _capi.rtg_tuning()
    check(lib.rtg_get_tuning(C.byref(t)), "rtg_get_tuning")
    return _tuning_dict(t)


def set_tuning(tile_w: int = 8, grab: int = 1, grid_pct: int = 100) -> None:
    """Install a launch tuning (rtg_set_tuning); pixels do not depend on it."""
    lib = _capi.load()
    t = _capi.rtg_tuning(tile_w=tile_w, grab=grab, grid_pct=grid_pct)
    check(lib.rtg_set_tuning(C.byref(t)), "rtg_set_tuning")


def autotune(items, reps: int = 3):
    """rtg_autotune over [(DeviceScene, rtg_params), ...] (one item: the frame
    path; several: one batch).  Installs the fastest tuning and returns
    (best {tile_w, grab, grid_pct}, [(tuning dict, median device ms), ...])."""
    lib = _capi.load()
    n = len(items)
    reqs = (_capi.rtg_request * max(1, n))()
    for k, (ds, p) in enumerate(items):
        reqs[k].scene = ds.h
        reqs[k].params = p
    best = _capi.rtg_tuning()
    cap = 64
    trials = (_capi.rtg_tune_trial * cap)()
    nt = C.c_int()
    check(lib.rtg_autotune(reqs, n, reps, C.byref(best), trials, cap, C.byref(nt)), "rtg_autotune")
    return _tuning_dict(best), [(_tuning_dict(trials[i].tuning), trials[i].device_ms)
                                for i in range(min(nt.value, cap))]


def write_png(path: str, rgb: np.ndarray, bottom_up: bool = True) -> None:
    """rtg_host_write_png: an H x W x 3 uint8 frame to an RGB PNG; with
    bottom_up (the reference's framebuffer order) row 0 becomes the image's
    bottom row."""
    a = np.ascontiguousarray(rgb, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("rgb must be H x W x 3")
    rc = _capi.load().rtg_host_write_png(path.encode(), a.ctypes.data_as(C.c_void_p), a.shape[1], a.shape[0], 0,
                                          int(bottom_up))
    if rc != 0:
        raise OSError(f"rtg_host_write_png({path!r}) failed ({rc})")


def gen_requests(n: int, rate: float = 0.0, seed: int = 3):
    """Request stream of workloads 3/4: [(scene index, width, height, arrival_s)]."""
    lib = _capi.load()
    buf = (_capi.rtg_gen_request * max(1, n))()
    check(lib.rtg_host_gen_requests(seed, n, rate, buf), "rtg_host_gen_requests")
    return [(buf[i].scene, buf[i].width, buf[i].height, buf[i].arrival_s) for i in range(n)]


def serve(items, mode: int = _capi.RTG_SERVE_BATCHED, slots: int = 0, max_batch: int = 0, hbm_budget: int = 0,
          graph_after: int = 0):
    """rtg_serve over [(DeviceScene, rtg_params, arrival_s), ...].
    Returns ([rgb per request], [(done_s, batch) per request], rtg_serve_stats).
    The result arrays are written once before the loop starts, as a server's
    pooled response buffers would be: the loop's latencies then hold no
    first-touch page faults of fresh host memory."""
    lib = _capi.load()
    n = len(items)
    reqs = (_capi.rtg_serve_request * max(1, n))()
    outs = []
    for k, (ds, p, t) in enumerate(items):
        rgb = np.empty((p.height, p.width, 3), np.uint8)
        rgb.fill(0)
        reqs[k].scene = ds.h
        reqs[k].params = p
        reqs[k].rgb = rgb.ctypes.data_as(C.POINTER(C.c_uint8))
        reqs[k].arrival_s = t
        outs.append(rgb)
    o = _capi.rtg_serve_opts(mode=mode, slots=slots, max_batch=max_batch, hbm_budget=hbm_budget,
                             graph_after=graph_after)
    st = _capi.rtg_serve_stats()
    check(lib.rtg_serve(reqs, n, C.byref(o), C.byref(st)), "rtg_serve")
    return outs, [(reqs[k].done_s, reqs[k].batch) for k in range(n)], st

