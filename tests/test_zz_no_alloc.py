"""rtg_render_device_noalloc (no counterpart in the reference): the frame of
rtg_render_device with the megakernel, from the library's fixed scratch array
instead of allocated workspace.  Bit-identical RGB, ids, hit distances and ray
counts on two calls in a row, the second smaller; sizes beyond the maximum
and the wavefront engine are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers from the HIP runtime librtg loaded."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(n)) == 0
        self.bufs.append(p)
        return p

    def fetch(self, p, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), 2) == 0
        return a

    def free_all(self):
        for p in self.bufs:
            assert self.lib.hipFree(p) == 0
        self.bufs = []


def _frame(rt, hip, ds, p, fn):
    W, H = p.width, p.height
    rgb, ids, t, cnt = hip.alloc(W * H * 3), hip.alloc(W * H * 8), hip.alloc(W * H * 4), hip.alloc(32)
    rc = fn(ds.h, C.byref(p), rgb, ids, t, cnt, None)
    assert rc == 0, rt._capi.last_error()
    return (hip.fetch(rgb, (H, W, 3), np.uint8), hip.fetch(ids, (H, W, 2), np.int32),
            hip.fetch(t, (H, W), np.float32), hip.fetch(cnt, (4,), np.uint64))


def test_no_alloc_matches_render_device(rt_gpu):
    rt = rt_gpu
    lib = rt._capi.load()
    hip = _Hip()
    cases = [(rt.HostScene.workload(1), rt.params(640, 480, max_level=3, engine=rt.RTG_ENGINE_MEGAKERNEL)),
             (rt.HostScene.workload(2, width=320, height=180, detail=20),
              rt.params(320, 180, max_level=4, spp=4, seed=9))]
    for k, (hs, p) in enumerate(cases):
        log(f"case: no-allocation call {k + 1}, {p.width}x{p.height}, depth {p.max_level}, spp {p.spp}")
        ds = rt.DeviceScene(hs)
        pm = rt.params(p.width, p.height, max_level=p.max_level, spp=p.spp, seed=p.seed,
                       engine=rt.RTG_ENGINE_MEGAKERNEL)
        ref = _frame(rt, hip, ds, pm, lib.rtg_render_device)
        got = _frame(rt, hip, ds, p, lib.rtg_render_device_noalloc)
        for a, b, what in zip(got, ref, ("rgb", "ids", "t", "ray counts")):
            assert np.array_equal(a, b), f"call {k + 1}: {what} differ from rtg_render_device"
        assert ref[3].sum() > 0
    # no stats pointer: the counts stay in the scratch
    hs, p = cases[1]
    ds = rt.DeviceScene(hs)
    rgb = hip.alloc(p.width * p.height * 3)
    assert lib.rtg_render_device_noalloc(ds.h, C.byref(p), rgb, None, None, None, None) == 0
    assert np.array_equal(hip.fetch(rgb, (p.height, p.width, 3), np.uint8), ref[0])
    # refused: sizes beyond the maximum, the wavefront engine, the FP64 re-test
    big = rt.params(16385, 64, max_level=1)
    assert lib.rtg_render_device_noalloc(ds.h, C.byref(big), rgb, None, None, None, None) == -6
    wf = rt.params(p.width, p.height, max_level=1, engine=rt.RTG_ENGINE_WAVEFRONT)
    assert lib.rtg_render_device_noalloc(ds.h, C.byref(wf), rgb, None, None, None, None) == -1
    r64 = rt.params(p.width, p.height, max_level=1, flags=rt.RTG_FLAG_FP64_RETEST)
    assert lib.rtg_render_device_noalloc(ds.h, C.byref(r64), rgb, None, None, None, None) == -1
    hip.free_all()
