"""ctypes bindings of librtg (include/rtg.h, include/rtg_host.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``raytrace_amd/librtg.so``.  Loading fails loudly when it is missing: there is
no Python or CPU fallback for the render path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtg.so")

RTG_OK = 0
RTG_E_INVALID = -1
RTG_E_DEVICE = -2
RTG_E_NOMEM = -3
RTG_E_CANCELLED = -4
RTG_E_STATE = -5
RTG_E_LIMIT = -6
RTG_E_PENDING = 1

RTG_MODE_RAYTRACE = 0x80
RTG_ENGINE_AUTO = 0
RTG_ENGINE_MEGAKERNEL = 1
RTG_ENGINE_WAVEFRONT = 2
RTG_FLAG_FP64_RETEST = 0x1
RTG_FLAG_KERNEL_TIMES = 0x2
RTG_FLAG_EAGER = 0x4
RTG_FLAG_ONE_STREAM = 0x8
RTG_SERVE_BATCHED = 0
RTG_SERVE_PLAIN = 1


class rtg_vec4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class rtg_material(C.Structure):
    _fields_ = [
        ("ambient", rtg_vec4), ("diffuse", rtg_vec4), ("specular", rtg_vec4), ("emission", rtg_vec4),
        ("shiness", C.c_float), ("reflect", C.c_float), ("refract", C.c_float), ("rfr", C.c_float),
        ("tex", C.c_int32), ("pad", C.c_int32 * 3),
    ]


class rtg_texture(C.Structure):
    _fields_ = [("w", C.c_int32), ("h", C.c_int32), ("bgr", C.POINTER(C.c_uint8))]


class rtg_object(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("show", C.c_int32), ("mtl", C.c_int32), ("tex", C.c_int32),
        ("position", rtg_vec4), ("radius", C.c_float), ("radius_sqr", C.c_float),
        ("part_begin", C.c_int32), ("part_count", C.c_int32),
        ("normal", rtg_vec4), ("axisx", rtg_vec4), ("axisy", rtg_vec4), ("vmin", rtg_vec4), ("vmax", rtg_vec4),
    ]


class rtg_triangle(C.Structure):
    _fields_ = [("p", rtg_vec4 * 3), ("n", rtg_vec4 * 3), ("uv", C.c_float * 6), ("pad", C.c_float * 2)]


class rtg_part(C.Structure):
    _fields_ = [
        ("tris", C.POINTER(rtg_triangle)), ("tri_count", C.c_int32), ("mtl", C.c_int32),
        ("bmin", rtg_vec4), ("bmax", rtg_vec4),
    ]


class rtg_light(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("on", C.c_int32), ("pad", C.c_int32 * 2),
        ("position", rtg_vec4), ("ambient", rtg_vec4), ("diffuse", rtg_vec4), ("specular", rtg_vec4),
        ("attenuation", rtg_vec4),
    ]


class rtg_camera(C.Structure):
    _fields_ = [
        ("position", rtg_vec4), ("u", rtg_vec4), ("v", rtg_vec4), ("n", rtg_vec4),
        ("width", C.c_int32), ("height", C.c_int32), ("fovy", C.c_float), ("zNear", C.c_float),
        ("zFar", C.c_float), ("pad", C.c_float * 3),
    ]


class rtg_scene_desc(C.Structure):
    _fields_ = [
        ("objects", C.POINTER(rtg_object)), ("n_objects", C.c_int32), ("n_parts", C.c_int32),
        ("parts", C.POINTER(rtg_part)), ("n_tris", C.c_int64), ("materials", C.POINTER(rtg_material)),
        ("n_materials", C.c_int32), ("n_textures", C.c_int32), ("textures", C.POINTER(rtg_texture)),
        ("lights", C.POINTER(rtg_light)), ("n_lights", C.c_int32), ("pad", C.c_int32),
        ("env_light", rtg_vec4), ("camera", rtg_camera),
    ]


def desc_of(ptr: int) -> rtg_scene_desc:
    """View of a flattened scene descriptor (valid while its HostScene lives)."""
    return rtg_scene_desc.from_address(ptr)


class rtg_params(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("max_level", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
        ("spp", C.c_int32), ("tile_compat", C.c_int32), ("engine", C.c_int32), ("flags", C.c_int32),
        ("seed", C.c_uint64),
    ]


class rtg_serve_request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", rtg_params), ("rgb", C.POINTER(C.c_uint8)),
                ("arrival_s", C.c_double), ("done_s", C.c_double), ("status", C.c_int32), ("batch", C.c_int32)]


class rtg_serve_opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("slots", C.c_int32), ("max_batch", C.c_int32), ("graph_after", C.c_int32),
                ("hbm_budget", C.c_uint64)]


class rtg_serve_stats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("mrays_per_s", C.c_double), ("req_per_s", C.c_double),
                ("rays_total", C.c_uint64), ("batches", C.c_int32), ("max_in_flight", C.c_int32)]


class rtg_gen_request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("pad", C.c_int32),
                ("arrival_s", C.c_double)]


class rtg_host_ray(C.Structure):
    _fields_ = [("origin", C.c_float * 4), ("direction", C.c_float * 4), ("type", C.c_int32), ("inside", C.c_int32)]


class rtg_host_hit(C.Structure):
    _fields_ = [("position", C.c_float * 4), ("normal", C.c_float * 4), ("u", C.c_float), ("v", C.c_float),
                ("distance", C.c_float), ("rfr", C.c_float), ("obj", C.c_int64), ("object", C.c_int64),
                ("inside", C.c_int32), ("hit", C.c_int32), ("has_tex", C.c_int32), ("pad", C.c_int32)]


class rtg_kernel_time(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("launches", C.c_uint32), ("pad", C.c_uint32), ("items", C.c_uint64),
                ("ms", C.c_double)]


class rtg_stats(C.Structure):
    _fields_ = [
        ("rays_primary", C.c_uint64), ("rays_shadow", C.c_uint64), ("rays_reflect", C.c_uint64),
        ("rays_refract", C.c_uint64), ("rays_total", C.c_uint64), ("render_ms", C.c_double),
        ("build_ms", C.c_double), ("retests_fp64", C.c_uint64), ("graph_replays", C.c_uint64),
    ]


class rtg_gen_opts(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("width", C.c_int32), ("height", C.c_int32), ("detail", C.c_int32),
        ("n_spheres", C.c_int32), ("n_lights", C.c_int32), ("index", C.c_int32), ("max_level", C.c_int32),
        ("pad", C.c_int32),
    ]


class rtg_gen_info(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("max_level", C.c_int32), ("spp", C.c_int32),
        ("n_tris", C.c_int64), ("n_objects", C.c_int32), ("n_lights", C.c_int32),
    ]


class rtg_request(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p), ("params", rtg_params), ("rgb", C.POINTER(C.c_uint8)),
        ("ids", C.POINTER(C.c_int32)),
    ]


class rtg_scene_info(C.Structure):
    _fields_ = [("bytes", C.c_uint64), ("build_ms", C.c_double), ("n_nodes", C.c_int32), ("n_prims", C.c_int32),
                ("rebuilds", C.c_uint64), ("refits", C.c_uint64), ("shading_updates", C.c_uint64)]


class rtg_tuning(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("grab", C.c_int32), ("grid_pct", C.c_int32), ("reserved", C.c_int32)]


class rtg_tune_trial(C.Structure):
    _fields_ = [("tuning", rtg_tuning), ("device_ms", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float

# name -> (restype, argtypes); every symbol declared in include/rtg*.h
SIGNATURES = {
    # rtg.h
    "rtg_init": (_I, [_I]),
    "rtg_shutdown": (_I, []),
    "rtg_last_error": (C.c_char_p, []),
    "rtg_abi_version": (_I, []),
    "rtg_scene_create": (_I, [_P, C.POINTER(_P)]),
    "rtg_scene_update": (_I, [_P, _P]),
    "rtg_scene_rebuild": (_I, [_P, _P]),
    "rtg_scene_stats": (_I, [_P, C.POINTER(rtg_scene_info)]),
    "rtg_scene_destroy": (_I, [_P]),
    "rtg_scene_bytes": (_I, [_P, C.POINTER(C.c_uint64)]),
    "rtg_scene_build_ms": (_I, [_P, C.POINTER(C.c_double)]),
    "rtg_render": (_I, [_P, C.POINTER(rtg_params), _P, _P, _P, C.POINTER(rtg_stats)]),
    "rtg_render_device": (_I, [_P, C.POINTER(rtg_params), _P, _P, _P, _P, _P]),
    "rtg_render_device_noalloc": (_I, [_P, C.POINTER(rtg_params), _P, _P, _P, _P, _P]),
    "rtg_render_async": (_I, [_P, C.POINTER(rtg_params), _P, _P, _P, C.POINTER(_P)]),
    "rtg_poll": (_I, [_P]),
    "rtg_cancel": (_I, [_P]),
    "rtg_job_wait": (_I, [_P, C.POINTER(rtg_stats)]),
    "rtg_job_free": (_I, [_P]),
    "rtg_render_batch": (_I, [C.POINTER(rtg_request), _I, C.POINTER(rtg_stats)]),
    "rtg_kernel_times": (_I, [C.POINTER(rtg_kernel_time), _I, C.POINTER(_I)]),
    "rtg_get_tuning": (_I, [C.POINTER(rtg_tuning)]),
    "rtg_set_tuning": (_I, [C.POINTER(rtg_tuning)]),
    "rtg_autotune": (_I, [C.POINTER(rtg_request), _I, _I, C.POINTER(rtg_tuning), C.POINTER(rtg_tune_trial), _I,
                          C.POINTER(_I)]),
    "rtg_serve": (_I, [C.POINTER(rtg_serve_request), _I, C.POINTER(rtg_serve_opts), C.POINTER(rtg_serve_stats)]),
    # rtg_host.h
    "rtg_host_scene_config": (_I, [_I, C.POINTER(rtg_gen_opts), C.POINTER(_P), C.POINTER(rtg_gen_info)]),
    "rtg_host_scene_new": (_I, [C.POINTER(_P)]),
    "rtg_host_gen_requests": (_I, [C.c_uint64, _I, C.c_double, C.POINTER(rtg_gen_request)]),
    "rtg_host_write_png": (_I, [C.c_char_p, _P, _I, _I, _I, _I]),
    "rtg_host_scene_destroy": (_I, [_P]),
    "rtg_host_scene_desc": (_P, [_P]),
    "rtg_host_add_sphere": (_I, [_P, _F]),
    "rtg_host_add_plane": (_I, [_P]),
    "rtg_host_add_cube": (_I, [_P, _F]),
    "rtg_host_add_ballplane": (_I, [_P, _F]),
    "rtg_host_add_model_obj": (_I, [_P, C.c_char_p, C.c_char_p]),
    "rtg_host_add_light": (_I, [_P, _I, C.POINTER(_F), C.POINTER(_F)]),
    "rtg_host_move_pos": (_I, [_P, _I, _I, C.POINTER(_F)]),
    "rtg_host_switch": (_I, [_P, _I, _I, _I]),
    "rtg_host_set_material": (_I, [_P, _I, C.POINTER(rtg_material)]),
    "rtg_host_set_position": (_I, [_P, _I, C.POINTER(_F)]),
    "rtg_host_set_light_position": (_I, [_P, _I, C.POINTER(_F)]),
    "rtg_host_set_env_light": (_I, [_P, C.POINTER(_F)]),
    "rtg_host_resize": (_I, [_P, _I, _I]),
    "rtg_host_raytrace": (_I, [_P, _I, _I, _I, _P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "rtg_host_raytrace_stop": (_I, [_P, _I, _I, C.c_double, _P, C.POINTER(C.c_int)]),
    "rtg_host_intersect": (_I, [_P, _I, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p]),
    "rtg_host_chg_light_comp": (_I, [_P, _I, _I, C.POINTER(C.c_float)]),
    "rtg_host_recolour": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "rtg_host_chg_mtl_library": (_I, [_P, _I, _I]),
    "rtg_host_delete": (_I, [_P, _I, _I]),
    "rtg_host_light": (_I, [_P, _I, C.c_void_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load librtg.so (raises FileNotFoundError / OSError when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (the render path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().rtg_last_error().decode(errors="replace")


class RtgError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} failed ({code}): {last_error()}")
        self.code = code


def check(code: int, where: str) -> int:
    if code < 0:
        raise RtgError(code, where)
    return code


def floats(*v):
    return (C.c_float * len(v))(*v)
