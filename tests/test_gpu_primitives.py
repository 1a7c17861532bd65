"""The rest of the reference's primitives on the GPU, against the REF32 oracle:
Box (Basic3DObject.cpp:274-300) and BallPlane (486-552) built through
Scene::AddCube / AddBallPlane (Scene.cpp:111-121, 146-155), with refraction so
that BallPlane's inside path runs; and a model loaded from OBJ + MTL + BMP
files through Scene::AddModel (Model.cpp:14-312, Scene.cpp:123-131) with a
textured part.  Same bar as the other parity tests."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import compare_frames, log

pytestmark = pytest.mark.gpu

ENGINES = [1, 2]  # megakernel, wavefront


def _oracle(hs, p):
    from oracle import pyoracle as po
    return po.OracleScene(hs.desc_ptr).render(p, po.ORACLE_REF32, threads=16)


def _check(g, o):
    s = compare_frames(g, o)
    log(f"  {s}")
    assert s["ids_mismatch_outside_band"] == 0, s
    assert s["ids_match"] >= 0.9999, s
    assert s["rgb_over_1"] == 0, s
    st, ost = g["stats"], o["stats"]
    assert (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract) == (
        ost.rays_primary, ost.rays_shadow, ost.rays_reflect, ost.rays_refract)
    return s


def _box_ballplane_scene(rt):
    hs = rt.HostScene.new()
    hs.add_plane()
    c0 = hs.add_cube(2.0)  # AddCube: at (0, len/2, 0), brass
    hs.set_position(c0, (-3.5, 1.0, 1.0))
    c1 = hs.add_cube(1.5)
    hs.set_position(c1, (3.0, 2.5, -1.0))
    hs.set_material(c1, (0.1, 0.1, 0.1), (0.2, 0.3, 0.7), (0.8, 0.8, 0.8), 40.0, reflect=0.5)
    b = hs.add_ballplane(0.4)  # 16 spheres on the plane through the origin
    hs.set_material(b, (0.1, 0.1, 0.1), (0.2, 0.6, 0.2), (0.5, 0.5, 0.5), 20.0, reflect=0.2, refract=0.6, rfr=1.4)
    s = hs.add_sphere(1.0)
    hs.set_position(s, (0.5, 3.0, 2.0))
    hs.set_material(s, (0.1, 0.1, 0.1), (0.7, 0.2, 0.2), (0.6, 0.6, 0.6), 30.0, reflect=0.1, refract=0.7, rfr=1.5)
    hs.add_light(2, (0.15, 0.55, 0.3), (0, 0, 1, 256))
    hs.set_light_position(0, (4, 9, 6, 1))
    hs.add_light(2, (0.2, 0.5, 0.3), (1, 0, 0.01, 1))
    hs.set_light_position(1, (-6, 7, 3, 1))
    return hs


@pytest.mark.parametrize("engine", ENGINES)
def test_box_and_ballplane_scene(rt_gpu, engine):
    rt = rt_gpu
    log(f"case: AddCube x2 + refracting AddBallPlane + sphere + plane, 640x480 depth 5, engine {engine}")
    hs = _box_ballplane_scene(rt)
    p = rt.params(640, 480, max_level=5, engine=engine)
    g = rt.DeviceScene(hs).render(p)
    o = _oracle(hs, p)
    s = _check(g, o)
    ids = o["ids"][..., 0]
    # every kind of object is on screen: plane 0, cubes 1-2, ballplane 3, sphere 4
    for k in range(5):
        assert (ids == k).sum() > 200, (k, int((ids == k).sum()))
    assert o["stats"].rays_refract > 1000
    assert s["rgb_max"] == 0 and s["t_exact"] == 1.0


def _write_asset(d):
    """A textured quad (part 'paint', map_Kd -> 8x4 BMP) and a two-triangle
    wedge (part 'steel', untextured, reflective), OBJ + MTL + BMP."""
    import struct
    rng = np.random.default_rng(9)
    w, h = 8, 4
    pix = rng.integers(0, 256, w * h * 3, dtype=np.uint8).tobytes()
    fh = b"BM" + struct.pack("<IHHI", 54 + len(pix), 0, 0, 54)
    ih = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, 0, 2835, 2835, 0, 0)
    (d / "paint.bmp").write_bytes(fh + ih + pix)
    (d / "asset.mtl").write_text("newmtl paint\nKa 0.2 0.2 0.2\nKd 0.9 0.8 0.7\nKs 0.3 0.3 0.3\nNs 12\n"
                                 "map_Kd textures\\paint.png\n"
                                 "newmtl steel\nKd 0.5 0.5 0.6\nKs 0.9 0.9 0.9\nNs 80\n")
    (d / "asset.obj").write_text(
        "v -1 0 0\nv 1 0 0\nv 1 1.2 0\nv -1 1.2 0\nv 0.4 0 0.6\nv 1 0 1\nv 1 0.8 0.6\n"
        "vn 0 0 1\nvn 0.6 0 0.8\nvn 0.3 0.4 0.87\n"
        "vt 0 0\nvt 2.5 0\nvt 2.5 1.5\nvt 0 1.5\n"
        "usemtl paint\nf 1/1/1 2/2/1 3/3/1 4/4/1\n"
        "usemtl steel\nf 5//2 6//2 7//3\nf 2//2 6//2 7//3\n")


@pytest.mark.parametrize("engine", ENGINES)
def test_loaded_obj_mtl_bmp_asset(rt_gpu, engine, tmp_path):
    rt = rt_gpu
    log(f"case: OBJ + MTL + BMP asset through Scene::AddModel, 640x480 depth 3, engine {engine}")
    _write_asset(tmp_path)
    cwd = os.getcwd()
    os.chdir(tmp_path)  # the reference opens the texture by base name from the working directory
    try:
        hs = rt.HostScene.new()
        hs.add_plane()
        m = hs.add_model_obj("asset.obj", "asset.mtl")
    finally:
        os.chdir(cwd)
    hs.move_pos(2, m, (-2.0, 0.0, -2.0))  # MY_MODEL_OBJECT: translate the model
    hs.add_light(2, (0.15, 0.55, 0.3), (0, 0, 1, 256))
    hs.set_light_position(0, (3, 10, 8, 1))
    from raytrace_amd import _capi
    d = _capi.desc_of(hs.desc_ptr)
    o_m = d.objects[m]
    assert o_m.part_count == 2 and d.n_textures >= 2  # the plane's checker and the BMP
    mt = d.materials[d.parts[o_m.part_begin].mtl]
    assert mt.tex >= 0 and (d.textures[mt.tex].w, d.textures[mt.tex].h) == (8, 4)
    p = rt.params(640, 480, max_level=3, engine=engine)
    g = rt.DeviceScene(hs).render(p)
    o = _oracle(hs, p)
    s = _check(g, o)
    ids = o["ids"]
    on_model = ids[..., 0] == m
    assert on_model.sum() > 2000, int(on_model.sum())
    # both parts are visible: triangles 0-1 the textured quad, 2-3 the wedge
    assert (on_model & (ids[..., 1] <= 1)).sum() > 1000 and (on_model & (ids[..., 1] >= 2)).sum() > 200
    # the texture shows: the quad's pixels take many colours
    quad = g["rgb"][on_model & (ids[..., 1] <= 1)]
    assert len(np.unique(quad.reshape(-1, 3), axis=0)) > 16
    # texture test mode (RTtex): the texel colour itself
    pt = rt.params(640, 480, max_level=0, mode=0x4, engine=engine)
    gt = rt.DeviceScene(hs).render(pt)
    ot = _oracle(hs, pt)
    _check(gt, ot)
    assert s["rgb_max"] <= 1


def _spheres_scene(rt, centres, radius=0.6, reflect=0.3):
    hs = rt.HostScene.new()
    for i, c in enumerate(centres):
        s = hs.add_sphere(radius)
        hs.set_position(s, tuple(float(x) for x in c))
        hs.set_material(s, (0.1, 0.1, 0.1), (0.3 + 0.05 * (i % 8), 0.4, 0.6 - 0.05 * (i % 8)), (0.5, 0.5, 0.5), 24.0,
                        reflect=reflect)
    hs.add_light(2, (0.2, 0.5, 0.3), (1, 0, 0.01, 1))
    hs.set_light_position(0, (3, 8, 6, 1))
    return hs


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_small_trees(rt_gpu, n):
    """Scenes of 1..9 spheres: a lone leaf (n <= 2) and the smallest 4-wide
    trees, nodes of 2, 3 and 4 children, leaves of one and two primitives."""
    rt = rt_gpu
    log(f"case: {n} spheres, 320x240 depth 3, both engines")
    rng = np.random.default_rng(100 + n)
    hs = _spheres_scene(rt, rng.uniform(-3, 3, (n, 3)))
    p = rt.params(320, 240, max_level=3)
    o = _oracle(hs, p)
    ds = rt.DeviceScene(hs)
    for engine in ENGINES:
        p.engine = engine
        s = _check(ds.render(p), o)
        assert s["rgb_max"] == 0 and s["t_exact"] == 1.0
    assert (o["ids"][..., 0] >= 0).sum() > 500


def test_coincident_primitives(rt_gpu):
    """Many spheres at one point share one Morton code (the radix tree splits
    them by index) and tie on every hit: the lowest object index wins
    (RayTracer.cpp:455-465).  Plus a row of identical spheres far from the
    others, so one node's children span very different extents."""
    rt = rt_gpu
    log("case: 12 coincident spheres + a distant row, 320x240 depth 2, both engines")
    centres = [(0.0, 1.0, 0.0)] * 12 + [(40.0 + 0.01 * k, -2.0, -30.0) for k in range(6)]
    hs = _spheres_scene(rt, np.array(centres), radius=0.9, reflect=0.2)
    p = rt.params(320, 240, max_level=2)
    o = _oracle(hs, p)
    ds = rt.DeviceScene(hs)
    for engine in ENGINES:
        p.engine = engine
        g = ds.render(p)
        s = _check(g, o)
        assert s["rgb_max"] == 0
    ids = o["ids"][..., 0]
    assert (ids == 0).sum() > 500 and not np.isin(ids, np.arange(1, 12)).any()
