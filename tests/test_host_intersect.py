"""DrawObject::intersect of the drop-in classes (3DElement.h:201), one ray
against one object on the host, against the oracle's primitive entry points
on random rays: the same distances bit for bit, and the hit records the
reference builds (normal, texture coordinates, identity, inside flag, IOR).
No GPU needed."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import raytrace_amd as rt
from raytrace_amd import _capi

BASE, SHADOW, REFLECT, REFRACT = 1, 2, 3, 4
MISS = np.float32(1e20)


@pytest.fixture(scope="module")
def lib():
    from oracle import pyoracle as po
    return po.load()


def f4(*v):
    v = [float(x) for x in v] + [0.0] * (4 - len(v))
    return (C.c_float * 4)(*v[:4])


def hit(hs, num, o, d, type_=BASE, inside=0, prev_obj=0, prev_dist=1e20, amin=0.0):
    r = _capi.rtg_host_ray(f4(*o), f4(*d), type_, inside)
    out = _capi.rtg_host_hit()
    _capi.check(_capi.load().rtg_host_intersect(hs.h, num, C.byref(r), prev_obj, prev_dist, amin, C.byref(out)),
                "rtg_host_intersect")
    return out


def obj(hs, k):
    return _capi.desc_of(hs.desc_ptr).objects[k]


def v3(q):
    return np.array([q.x, q.y, q.z], np.float32)


def unit(rng):
    d = rng.standard_normal(3)
    return (d / np.linalg.norm(d)).astype(np.float32)


def test_sphere(lib):
    hs = rt.HostScene.new()
    k = hs.add_sphere(1.3)
    hs.set_position(k, (0.5, 2.0, -1.0))
    hs.set_material(k, (0.1,) * 3, (0.5,) * 3, (0.5,) * 3, 10.0, refract=0.5, rfr=1.45)
    c = v3(obj(hs, k).position)
    rng = np.random.default_rng(1)
    ti = C.c_float()
    n_hit = 0
    for _ in range(3000):
        o = (c + unit(rng) * rng.uniform(1.5, 10)).astype(np.float32)
        d = unit(rng) if rng.random() < 0.5 else ((c + rng.uniform(-1, 1, 3) - o) / np.linalg.norm(
            c + rng.uniform(-1, 1, 3) - o)).astype(np.float32)
        t = lib.oracle_sphere_test(f4(*o), f4(*d), f4(*c), np.float32(1.3), C.byref(ti))
        h = hit(hs, k, o, d)
        if t < MISS and t > 1e-6:
            n_hit += 1
            assert h.hit and h.distance == t, (h.distance, t)
            P = o + d * np.float32(t)
            nrm = (P - c) / np.linalg.norm(P - c)
            assert np.allclose(list(h.normal)[:3], nrm, atol=1e-5)
            assert h.obj == h.object and h.inside == 0xFF and h.rfr == np.float32(1.45)
            # the ray continues from the hit, inside the sphere (the refraction
            # child): the far side with the inward normal, IOR 1 on leaving
            o2 = np.array(list(h.position)[:3], np.float32)
            h2 = hit(hs, k, o2, d, REFRACT, 0xFF, h.obj)
            lib.oracle_sphere_test(f4(*o2), f4(*d), f4(*c), np.float32(1.3), C.byref(ti))
            if not ti.value > 1e-6:  # a grazing hit: no chord (NaN or ~0) and no exit
                assert not h2.hit
                continue
            assert h2.hit and h2.distance == ti.value and h2.rfr == 1.0 and h2.inside == 0
            P2 = o2 + d * np.float32(ti.value)
            assert np.allclose(list(h2.normal)[:3], (c - P2) / np.linalg.norm(c - P2), atol=1e-5)
            # a reflection ray inside keeps the material's IOR; a shadow ray is blocked at the radius
            assert hit(hs, k, o2, d, REFLECT, 0xFF, h.obj).rfr == np.float32(1.45)
            hs_ = hit(hs, k, o2, d, SHADOW, 0xFF, h.obj)
            assert hs_.distance == np.float32(1.3)
            # from outside, the sphere the ray leaves is skipped
            assert not hit(hs, k, o2, d, REFLECT, 0, h.obj).hit
            # a nearer previous hit stays
            assert not hit(hs, k, o, d, prev_dist=t * 0.5).hit
        else:
            assert not h.hit
    assert n_hit > 500


def test_plane(lib):
    hs = rt.HostScene.new()
    k = hs.add_plane()
    hs.move_pos(2, k, (3, -4, 2))  # Plane::rotate: a tilted plane
    ob = obj(hs, k)
    p0, n, ax, ay = v3(ob.position), v3(ob.normal), v3(ob.axisx), v3(ob.axisy)
    rng = np.random.default_rng(2)
    uv = (C.c_float * 2)()
    n_hit = 0
    for _ in range(3000):
        o = (rng.uniform(-10, 10, 3)).astype(np.float32)
        d = unit(rng)
        t = lib.oracle_plane_hit(f4(*o), f4(*d), f4(*p0), f4(*n), f4(*ax), f4(*ay), uv)
        h = hit(hs, k, o, d)
        if t < MISS:
            n_hit += 1
            assert h.hit and h.distance == t and (h.u, h.v) == (uv[0], uv[1]), (h.distance, t, h.u, uv[0])
            assert list(h.normal)[:3] == list(n) and h.has_tex and h.obj == h.object
            assert not hit(hs, k, o, d, prev_obj=h.object).hit  # its own reflection ray skips it
        else:
            assert not h.hit
    assert n_hit > 1000


def test_box(lib):
    hs = rt.HostScene.new()
    k = hs.add_cube(1.5)  # Box(1.5): [-0.75, 0.75]^3 at (0, 0.75, 0)
    hs.set_position(k, (1.0, 0.5, -2.0))
    ob = obj(hs, k)
    pos, mn, mx = v3(ob.position), v3(ob.vmin), v3(ob.vmax)
    assert list(mn) == [-0.75] * 3 and list(mx) == [0.75] * 3
    rng = np.random.default_rng(3)
    nrm = (C.c_float * 4)()
    n_hit = 0
    for _ in range(3000):
        o = (pos + unit(rng) * rng.uniform(1.5, 8)).astype(np.float32)
        tgt = pos + rng.uniform(-1, 1, 3)
        d = ((tgt - o) / np.linalg.norm(tgt - o)).astype(np.float32)
        t = lib.oracle_box_hit(f4(*o), f4(*d), f4(*pos), f4(*mn), f4(*mx), nrm)
        h = hit(hs, k, o, d)
        if t < MISS and t > 1e-6:
            n_hit += 1
            assert h.hit and h.distance == t
            assert np.array_equal(np.array(list(h.normal)[:3]), np.array(list(nrm)[:3]), equal_nan=True)
            assert not hit(hs, k, o, d, prev_obj=h.object).hit
        else:
            assert not h.hit
    assert n_hit > 1000


def test_ballplane(lib):
    hs = rt.HostScene.new()
    k = hs.add_ballplane(0.5)
    hs.set_material(k, (0.1,) * 3, (0.5,) * 3, (0.5,) * 3, 10.0, refract=0.5, rfr=1.3)
    ob = obj(hs, k)
    pos, ax, ay, n = v3(ob.position), v3(ob.axisx), v3(ob.axisy), v3(ob.normal)
    r = np.float32(0.5)
    rng = np.random.default_rng(4)
    oi = (C.c_int32 * 2)()
    of = (C.c_float * 4)()
    n_hit = 0
    for _ in range(2000):
        tgt = pos + ax * rng.uniform(-4, 4) + ay * rng.uniform(-4, 4)
        o = (tgt + unit(rng) * rng.uniform(2, 8)).astype(np.float32)
        d = ((tgt - o) / np.linalg.norm(tgt - o)).astype(np.float32)
        t = lib.oracle_ballplane_hit(f4(*o), f4(*d), f4(*pos), f4(*ax), f4(*ay), r, np.float32(1.3), -1, 0, BASE, oi,
                                     of)
        h = hit(hs, k, o, d)
        if t < MISS:
            n_hit += 1
            assert h.hit and h.distance == t
            assert h.obj - h.object == oi[0] + 1  # identity this + k, k = 1..16 in loop order
            assert list(h.normal)[:3] == list(of)[:3] and h.rfr == of[3]
            # inside sphere k: the exit, then the shadow block at the radius
            o2 = np.array(list(h.position)[:3], np.float32)
            t2 = lib.oracle_ballplane_hit(f4(*o2), f4(*d), f4(*pos), f4(*ax), f4(*ay), r, np.float32(1.3), oi[0], 1,
                                          REFRACT, oi, of)
            h2 = hit(hs, k, o2, d, REFRACT, 0xFF, h.obj)
            if t2 < MISS:
                assert h2.distance == t2 and h2.obj - h2.object == oi[0] + 1 and h2.rfr == of[3]
            else:
                assert not h2.hit
            assert hit(hs, k, o2, d, SHADOW, 0xFF, h.obj).distance == r
        else:
            assert not h.hit
    assert n_hit > 300


def test_model_against_brute_force(lib):
    """Model::intersect over RTPrepare's octant bins equals the nearest
    triangle of a brute-force scan (TriangleTest on every triangle), except
    where the octant culling grazes; the normal and uv are the barycentric
    blends; the hit copy is skipped by the ray leaving it; with `min` the scan
    stops at the first triangle closer than it."""
    hs = rt.HostScene.workload(3, index=1, detail=3000)  # a height field
    d_ = _capi.desc_of(hs.desc_ptr)
    k = [i for i in range(d_.n_objects) if d_.objects[i].type == 3][0]
    ob = d_.objects[k]
    pos = v3(ob.position)
    tris = []
    for pi in range(ob.part_begin, ob.part_begin + ob.part_count):
        part = d_.parts[pi]
        for t in range(part.tri_count):
            tr = part.tris[t]
            p = [v3(tr.p[q]) for q in range(3)]
            tris.append((p[1] - p[0], p[2] - p[0], p[0] + pos, [v3(tr.n[q]) for q in range(3)]))
    lo, hi = v3(ob.vmin) + pos, v3(ob.vmax) + pos
    rng = np.random.default_rng(5)
    bary = (C.c_float * 3)()
    n_hit = agree = 0
    for _ in range(150):
        tgt = lo + (hi - lo) * rng.uniform(0.1, 0.9, 3)
        o = (tgt + unit(rng) * 12 + np.array([0, 10, 0])).astype(np.float32)
        d = ((tgt - o) / np.linalg.norm(tgt - o)).astype(np.float32)
        best, bn = MISS, None
        for e1, e2, p0, ns in tris:
            t = lib.oracle_tri_test(f4(*o), f4(*d), f4(*e1), f4(*e2), f4(*p0), bary)
            if t < best:
                best, bn = t, (list(bary), ns)
        h = hit(hs, k, o, d)
        if best < MISS:
            n_hit += 1
            if h.hit and h.distance == best:
                agree += 1
                w = np.array(bn[0], np.float32)
                nb = bn[1][0] * w[0] + bn[1][1] * w[1] + bn[1][2] * w[2]
                assert np.allclose(list(h.normal)[:3], nb / np.linalg.norm(nb), atol=1e-5)
                # continuing from the hit, the copy it hit is skipped
                o2 = np.array(list(h.position)[:3], np.float32)
                h2 = hit(hs, k, o2, d, REFLECT, 0, h.obj)
                assert not h2.hit or h2.obj != h.obj
                # any-hit: with min above the nearest, the first closer copy ends the scan
                ha = hit(hs, k, o, d, SHADOW, amin=1e19)
                assert ha.hit and ha.distance >= best
        else:
            assert not h.hit or h.distance > 0
    assert n_hit > 60 and agree >= n_hit - 2, (n_hit, agree)
