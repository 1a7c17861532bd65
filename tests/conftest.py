"""Shared test setup: markers, import paths, helpers."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EMUL_PATH = os.path.join(ROOT, "tests", "_build", "libemul.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels of librtg)")


def log(msg: str) -> None:
    """Progress line that names the case before a GPU step."""
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


_emul = None


def emul_lib():
    """CPU emulation of the device code (tests/cpu_emul/emul.cpp)."""
    global _emul
    if _emul is None:
        if not os.path.exists(EMUL_PATH):
            raise FileNotFoundError(f"{EMUL_PATH} missing: run __graft_entry__.build()")
        lib = C.CDLL(EMUL_PATH)
        lib.emul_render.argtypes = [C.c_void_p] * 2 + [C.c_int, C.c_void_p, C.c_int] + [C.c_void_p] * 5
        lib.emul_render.restype = C.c_int
        lib.emul_max_stack.restype = C.c_int
        _emul = lib
    return _emul


def emul_render(desc_ptr: int, p, threads: int = 8, tiles=None) -> dict:
    W, H = p.width, p.height
    rgb = np.full((H, W, 3), 127, np.uint8)
    ids = np.full((H, W, 2), -1, np.int32)
    t = np.full((H, W), np.float32(1e20), np.float32)
    cnt = np.zeros(5, np.uint64)  # ray counts by type, FP64 re-tests
    depth = C.c_int()
    tl = np.ascontiguousarray(np.asarray(tiles, np.int32)) if tiles is not None else None
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    rc = emul_lib().emul_render(C.c_void_p(desc_ptr), C.byref(p), threads, ptr(tl), 0 if tl is None else len(tl),
                                ptr(rgb), ptr(ids), ptr(t), ptr(cnt), C.byref(depth))
    assert rc == 0
    return {"rgb": rgb, "ids": ids, "t": t, "counts": cnt[:4], "retests": int(cnt[4]), "depth": depth.value,
            "max_stack": emul_lib().emul_max_stack()}


def compare_frames(test: dict, ref: dict, region=None) -> dict:
    """Parity statistics of `test` against the REF32 oracle `ref`.

    ids: fraction of pixels whose primary (object, primitive) match; mismatches
    must lie in the oracle's near-tie / cut-edge band.  rgb: max channel
    difference where the ids match (tolerance 1/255 = one quantisation step).
    """
    sel = np.ones(ref["rgb"].shape[:2], bool) if region is None else region
    ids_eq = (test["ids"] == ref["ids"]).all(-1)
    band = (ref["flags"] & 3) != 0
    diff = np.abs(test["rgb"].astype(np.int16) - ref["rgb"].astype(np.int16)).max(-1)
    return {
        "n": int(sel.sum()),
        "ids_match": float(ids_eq[sel].mean()) if sel.any() else 1.0,
        "ids_mismatch_outside_band": int((~ids_eq & ~band & sel).sum()),
        "rgb_over_1": int(((diff > 1) & ids_eq & sel).sum()),
        "rgb_max": int(diff[sel & ids_eq].max()) if (sel & ids_eq).any() else 0,
        "t_exact": float((test["t"] == ref["t"])[sel & ids_eq].mean()) if (sel & ids_eq).any() else 1.0,
    }


def tile_mask(W, H, tiles):
    m = np.zeros((H, W), bool)
    tx = (W + 63) // 64
    for t in tiles:
        bx, by = t % tx, t // tx
        m[by * 64:min(H, by * 64 + 64), bx * 64:min(W, bx * 64 + 64)] = True
    return m


def sample_tiles(W, H, frac=16, seed=7):
    """Seeded 1/frac sample of the full-frame 64x64 tiles."""
    tx, ty = (W + 63) // 64, (H + 63) // 64
    rng = np.random.default_rng(seed)
    n = max(1, (tx * ty) // frac)
    return np.sort(rng.choice(tx * ty, size=n, replace=False)).astype(np.int32)


@pytest.fixture(scope="session")
def rt_gpu():
    import raytrace_amd as rt
    rt.init(0)
    return rt


def fp64_stats(test: dict, o64: dict, region=None) -> dict:
    """Pixel-error statistics of a frame against the FP64 oracle: primary-id
    mismatch fraction, RGB mean / max absolute error (in 1/255 steps) and the
    fraction of pixels off by more than one step."""
    sel = np.ones(o64["rgb"].shape[:2], bool) if region is None else region
    ids_eq = (test["ids"] == o64["ids"]).all(-1)
    d = np.abs(test["rgb"].astype(np.int16) - o64["rgb"].astype(np.int16))
    return {
        "n": int(sel.sum()),
        "ids_mismatch": float((~ids_eq & sel).sum() / max(1, sel.sum())),
        "rgb_mae": float(d[sel].mean()) if sel.any() else 0.0,
        "rgb_max": int(d[sel].max()) if sel.any() else 0,
        "rgb_over_1": float(((d.max(-1) > 1) & sel).sum() / max(1, sel.sum())),
    }
