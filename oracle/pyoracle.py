"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (raytrace_amd/)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

ORACLE_REF32 = 0
ORACLE_FP64 = 1
FLAG_NEAR_TIE = 0x1
FLAG_CUT_EDGE = 0x2


class oracle_stats(C.Structure):
    _fields_ = [
        ("rays_primary", C.c_uint64), ("rays_shadow", C.c_uint64), ("rays_reflect", C.c_uint64),
        ("rays_refract", C.c_uint64), ("pixels", C.c_uint64), ("seconds", C.c_double),
        ("threads", C.c_int32), ("pad", C.c_int32),
    ]

    def rays(self) -> int:
        return self.rays_primary + self.rays_shadow + self.rays_reflect + self.rays_refract


class oracle_traffic(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
        ("sph_tests", C.c_uint64), ("n_nodes", C.c_uint64), ("n_prims", C.c_uint64),
        ("shadow_rays", C.c_uint64), ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
        ("shadow_sph_tests", C.c_uint64), ("shadow_skipped", C.c_uint64),
    ]


_libs = {}
_F4 = C.c_float * 4


def load(path: str | None = None) -> C.CDLL:
    """Load the oracle library (default build, or another build at `path`)."""
    path = path or LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (make -C oracle)")
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.oracle_scene_create.argtypes = [P, C.POINTER(P)]
    lib.oracle_scene_create.restype = C.c_int
    lib.oracle_scene_destroy.argtypes = [P]
    lib.oracle_scene_destroy.restype = None
    lib.oracle_render.argtypes = [P, P, C.c_int, C.c_int, P, C.c_int, P, P, P, P, P, C.POINTER(oracle_stats)]
    lib.oracle_render.restype = C.c_int
    lib.oracle_traffic_count.argtypes = [P, P, C.c_int, P, C.c_int, C.POINTER(oracle_traffic)]
    lib.oracle_traffic_count.restype = C.c_int
    lib.oracle_dot.argtypes = [_F4, _F4]
    lib.oracle_dot.restype = C.c_float
    lib.oracle_cross.argtypes = [_F4, _F4, _F4]
    lib.oracle_normalize.argtypes = [_F4, _F4]
    lib.oracle_tri_test.argtypes = [_F4, _F4, _F4, _F4, _F4, C.c_float * 3]
    lib.oracle_tri_test.restype = C.c_float
    lib.oracle_sphere_test.argtypes = [_F4, _F4, _F4, C.c_float, C.POINTER(C.c_float)]
    lib.oracle_sphere_test.restype = C.c_float
    lib.oracle_plane_test.argtypes = [_F4, _F4, _F4, _F4]
    lib.oracle_plane_test.restype = C.c_float
    lib.oracle_border_test.argtypes = [_F4, _F4, _F4, _F4, C.POINTER(C.c_float)]
    lib.oracle_border_test.restype = C.c_float
    lib.oracle_border_test_ex.argtypes = [_F4, _F4, _F4, _F4, C.c_int32 * 8]
    lib.oracle_border_test_ex.restype = C.c_float
    lib.oracle_raygen.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, _F4, _F4]
    lib.oracle_quantize.argtypes = [C.c_float]
    lib.oracle_quantize.restype = C.c_uint8
    lib.oracle_plane_hit.argtypes = [_F4] * 6 + [C.c_float * 2]
    lib.oracle_plane_hit.restype = C.c_float
    lib.oracle_box_hit.argtypes = [_F4] * 6
    lib.oracle_box_hit.restype = C.c_float
    lib.oracle_ballplane_hit.argtypes = [_F4] * 5 + [C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                                     C.c_int32 * 2, _F4]
    lib.oracle_ballplane_hit.restype = C.c_float
    lib.oracle_texel.argtypes = [P, C.c_float, C.c_float, C.c_float * 3]
    lib.oracle_texel.restype = None
    _libs[path] = lib
    return lib


def f4(v) -> "C.Array":
    v = list(v) + [0.0] * (4 - len(v))
    return _F4(*[float(x) for x in v[:4]])


class OracleScene:
    """Oracle preparation (Model::RTPrepare + canonical BVH) of a scene descriptor."""

    def __init__(self, desc_ptr: int, lib: C.CDLL | None = None):
        self.lib = lib or load()
        h = C.c_void_p()
        rc = self.lib.oracle_scene_create(C.c_void_p(desc_ptr), C.byref(h))
        if rc != 0:
            raise RuntimeError("oracle_scene_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.oracle_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, params, mode=ORACLE_REF32, threads=None, tiles=None, want=("rgb", "ids", "t", "rgbf", "flags")):
        """Render; returns dict of numpy arrays (rows bottom-up, as the reference stores them)."""
        W, H = params.width, params.height
        threads = threads or max(1, os.cpu_count() or 1)
        out = {"rgb": np.full((H, W, 3), 127, np.uint8)}
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        ids = np.full((H, W, 2), -1, np.int32) if "ids" in want else None
        t = np.full((H, W), np.float32(1e20), np.float32) if "t" in want else None
        rgbf = np.zeros((H, W, 3), np.float32) if "rgbf" in want else None
        flags = np.zeros((H, W), np.uint8) if "flags" in want else None
        tl = None
        nt = 0
        if tiles is not None:
            tl = np.ascontiguousarray(np.asarray(tiles, np.int32))
            nt = len(tl)
        st = oracle_stats()
        rc = self.lib.oracle_render(self.h, C.byref(params), mode, threads, ptr(tl), nt, ptr(out["rgb"]), ptr(ids),
                                    ptr(t), ptr(rgbf), ptr(flags), C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render failed")
        out.update(ids=ids, t=t, rgbf=rgbf, flags=flags, stats=st)
        return out

    def traffic(self, params, threads=None, tiles=None) -> oracle_traffic:
        threads = threads or max(1, os.cpu_count() or 1)
        tl = None
        nt = 0
        if tiles is not None:
            tl = np.ascontiguousarray(np.asarray(tiles, np.int32))
            nt = len(tl)
        tr = oracle_traffic()
        rc = self.lib.oracle_traffic_count(self.h, C.byref(params), threads,
                                           tl.ctypes.data_as(C.c_void_p) if tl is not None else None, nt,
                                           C.byref(tr))
        if rc != 0:
            raise RuntimeError("oracle_traffic_count failed")
        return tr
