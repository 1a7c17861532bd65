"""Device-resident frames (rtg_render_device, the caller's buffers): repeated
frames of one shape, with frames of another scene launched between them and a
material and light change in between, must each equal the host render of the
same scene state.  The buffers come from the HIP runtime librtg itself loaded,
so nothing else in the process initialises the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime librtg is linked against (already loaded): device
    buffers for rtg_render_device without another runtime in the process."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")  # matched by soname: the copy librtg loaded

    def alloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(n)) == 0
        return p

    def fetch(self, p, shape):
        a = np.empty(shape, np.uint8)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), 2) == 0  # device to host
        return a

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def _device_frame(rt, ds, p, rgb):
    lib = rt._capi.load()
    rc = lib.rtg_render_device(ds.h, C.byref(p), rgb, None, None, None, None)  # the library's stream
    assert rc == 0, rt._capi.last_error()


def test_repeated_device_frames(rt_gpu):
    rt = rt_gpu
    hip = _Hip()
    hs = rt.HostScene.workload(2, width=480, height=270)
    ds = rt.DeviceScene(hs)
    other = rt.DeviceScene(rt.HostScene.workload(1))
    p = rt.params(480, 270, max_level=4)
    po = rt.params(320, 240, max_level=3)
    rgb, rgb_o = hip.alloc(270 * 480 * 3), hip.alloc(240 * 320 * 3)
    ref = ds.render(p, ids=False, tprim=False)["rgb"]
    ref_o = other.render(po, ids=False, tprim=False)["rgb"]
    log("case: workload 2 at 480x270, 6 device frames with other frames between")
    for k in range(6):
        _device_frame(rt, ds, p, rgb)
        assert (hip.fetch(rgb, (270, 480, 3)) == ref).all(), f"device frame {k} differs from the host render"
        if k % 2:
            _device_frame(rt, other, po, rgb_o)
            assert (hip.fetch(rgb_o, (240, 320, 3)) == ref_o).all()
    # material and light changes re-upload the shading tables in place: the
    # replay must render the new state
    hs.set_material(1, (0.1, 0.1, 0.1), (0.3, 0.4, 0.5), (0.5, 0.5, 0.5), 20.0, reflect=0.2, refract=0.6, rfr=1.4)
    hs.set_light_position(0, (2.0, 9.0, 3.0, 1.0))
    ds.update(hs)
    ref2 = ds.render(p, ids=False, tprim=False)["rgb"]
    assert (ref2 != ref).any()
    for k in range(3):
        _device_frame(rt, ds, p, rgb)
        assert (hip.fetch(rgb, (270, 480, 3)) == ref2).all(), f"device frame {k} after the update differs"
    hip.free(rgb)
    hip.free(rgb_o)


def test_large_device_frame_on_a_caller_stream(rt_gpu):
    """A frame of 2^20 samples or more on a caller's stream, all levels
    enqueued at once, alternates its levels' shadow rays between the side
    stream and the arena's own stream: every frame must equal the render with
    all kernels on one stream."""
    rt = rt_gpu
    hip = _Hip()
    ds = rt.DeviceScene(rt.HostScene.workload(2))
    W, H = 1280, 1024  # 1.31 M samples
    ref = ds.render(rt.params(W, H, max_level=4, flags=rt.RTG_FLAG_ONE_STREAM), ids=False, tprim=False)["rgb"]
    p = rt.params(W, H, max_level=4)
    s = C.c_void_p()
    assert hip.lib.hipStreamCreate(C.byref(s)) == 0
    rgb = hip.alloc(W * H * 3)
    lib = rt._capi.load()
    log(f"case: workload 2 at {W}x{H}, depth 4, 3 device frames on a caller stream")
    for k in range(3):
        rc = lib.rtg_render_device(ds.h, C.byref(p), rgb, None, None, None, s)
        assert rc == 0, rt._capi.last_error()
        assert hip.lib.hipStreamSynchronize(s) == 0
        assert (hip.fetch(rgb, (H, W, 3)) == ref).all(), f"device frame {k} differs from the one-stream render"
    hip.free(rgb)
    assert hip.lib.hipStreamDestroy(s) == 0
