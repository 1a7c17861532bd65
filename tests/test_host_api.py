"""Host side of the drop-in API (raytrace_amd/host): the scene values the
reference's Scene / Light / Plane / Model code produces, checked against
closed forms.  No GPU needed."""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import pytest

import raytrace_amd as rt
from raytrace_amd import _capi


def desc(hs):
    return _capi.desc_of(hs.desc_ptr)


def v(a):
    return np.array([a.x, a.y, a.z], np.float64)


def test_default_plane_frame():
    """Plane() = rotate((0,36,0)) (Basic3DObject.cpp:304-330): normal ~ +y through
    the origin, axes orthonormal and in the plane, 4x4 grey checker texture."""
    hs = rt.HostScene.new()
    hs.add_plane()
    d = desc(hs)
    o = d.objects[0]
    assert o.type == 4
    n, ax, ay = v(o.normal), v(o.axisx), v(o.axisy)
    assert np.allclose(n, [0, 1, 0], atol=1e-7) and np.allclose(v(o.position), 0)
    assert abs(np.linalg.norm(ax) - 1) < 1e-6 and abs(np.linalg.norm(ay) - 1) < 1e-6
    assert abs(ax @ n) < 1e-6 and abs(ay @ n) < 1e-6 and abs(ax @ ay) < 1e-6
    t = d.textures[o.tex]
    tex = np.ctypeslib.as_array(t.bgr, shape=(t.w * t.h * 3,)).reshape(4, 4, 3)
    assert (t.w, t.h) == (4, 4)
    for a in range(4):
        for b in range(4):
            assert (tex[a, b] == (255 if (a & 2) == (b & 2) else 127)).all()
    assert d.materials[o.mtl].reflect == np.float32(0.6)  # AddPlane material


def test_add_light_normalises_and_scales():
    """AddLight (Scene.cpp:83-97): components divided by their sum, then
    SetLumi(atte.w) scales them; attenuation keeps (x, y, z)."""
    hs = rt.HostScene.new()
    hs.add_light(2, (0.15, 0.55, 0.3), (0, 0, 1, 256))
    L = desc(hs).lights[0]
    assert L.type == 2 and L.on == 1
    assert np.allclose([L.ambient.x, L.diffuse.x, L.specular.x], [0.15 * 256, 0.55 * 256, 0.3 * 256], rtol=1e-6)
    assert (L.attenuation.x, L.attenuation.y, L.attenuation.z, L.attenuation.w) == (0, 0, 1, 256)
    assert L.position.w == 1.0  # point light
    # default position: rangy 90, rangz 0, rdis 16 -> (0, ~0, 16)
    assert np.allclose(v(L.position), [0, 0, 16], atol=1e-5)


def test_light_move_spherical():
    """Light::move (3DElement.cpp:533-545) as main.cpp:166-167 drives it:
    angles (90 - 45, 0 + 45) at distance 16."""
    hs = rt.HostScene.new()
    hs.add_light(1, (0.1, 0.45, 0.45))
    hs.move_pos(1, 0, (-45, 45, 0))
    L = desc(hs).lights[0]
    s = math.sin(math.radians(45))
    assert np.allclose(v(L.position), [16 * s * s, 16 * math.cos(math.radians(45)), 16 * s * s], rtol=1e-6)
    assert L.position.w == 0.0  # parallel light


def test_light_cap_is_eight():
    hs = rt.HostScene.new()
    idx = [hs.add_light(2, (1, 1, 1)) for _ in range(9)]
    assert idx[:8] == list(range(8)) and idx[8] == 255


def test_add_sphere_pose_and_material():
    hs = rt.HostScene.new()
    i = hs.add_sphere(1.25)
    o = desc(hs).objects[i]
    assert o.type == 1 and (o.position.x, o.position.y, o.position.z) == (0, 1.25, 0)
    assert o.radius == np.float32(1.25) and o.radius_sqr == np.float32(1.25) * np.float32(1.25)
    m = desc(hs).materials[o.mtl]
    assert m.reflect == np.float32(0.35) and m.shiness == 100  # "bas-sphere"


def test_switch_hides_object():
    hs = rt.HostScene.new()
    hs.add_sphere(1.0)
    assert hs.switch(2, 0, False) == 1  # returns the previous state
    assert desc(hs).objects[0].show == 0
    assert hs.switch(2 | 0x80, 0, True) == 0  # toggle
    assert desc(hs).objects[0].show == 1


def test_workload2_icosphere_geometry():
    """Workload 2: 20 parts x F^2 triangles, extent-8 rescale, unit normals."""
    hs = rt.HostScene.workload(2, detail=10)
    d = desc(hs)
    assert hs.info.n_tris == 2000 and d.n_parts == 20
    o = d.objects[0]
    ext = [o.vmax.x - o.vmin.x, o.vmax.y - o.vmin.y, o.vmax.z - o.vmin.z]
    assert abs(max(ext) - 8.0) < 1e-5  # Model.cpp:185-196
    pts, nrm = [], []
    for p in range(d.n_parts):
        part = d.parts[p]
        assert part.tri_count == 100
        for t in range(part.tri_count):
            tri = part.tris[t]
            pts += [v(tri.p[k]) for k in range(3)]
            nrm += [v(tri.n[k]) for k in range(3)]
    pts, nrm = np.array(pts), np.array(nrm)
    r = np.linalg.norm(pts, axis=1)
    assert 3.0 < r.min() and r.max() < 4.5  # radius 1 +- 0.1 noise, scaled to an extent of 8
    lo, hi = v(o.vmin), v(o.vmax)
    assert (pts >= lo - 1e-6).all() and (pts <= hi + 1e-6).all()
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)
    assert ((nrm * pts).sum(1) > 0).mean() > 0.99  # outward


def test_obj_loader_quads_parts_rescale(tmp_path):
    """OBJ/MTL reading (Model.cpp:14-281): quads split along 0-2, one part per
    usemtl, material by name, extent-8 rescale, vn normals kept."""
    obj = tmp_path / "m.obj"
    mtl = tmp_path / "m.mtl"
    mtl.write_text("newmtl red\nKd 1 0 0\nKa 0.1 0 0\nNs 20\nnewmtl blue\nKd 0 0 1\n")
    obj.write_text("v 0 0 0\nv 2 0 0\nv 2 2 0\nv 0 2 0\nv 0 0 1\nvn 0 0 1\n"
                   "usemtl red\nf 1//1 2//1 3//1 4//1\nusemtl blue\nf 1//1 2//1 5//1\n")
    hs = rt.HostScene.new()
    i = hs.add_model_obj(str(obj), str(mtl))
    d = desc(hs)
    o = d.objects[i]
    assert o.type == 3 and o.part_count == 2
    p0, p1 = d.parts[o.part_begin], d.parts[o.part_begin + 1]
    assert (p0.tri_count, p1.tri_count) == (2, 1)
    # extent 2 -> scale 4
    t = p0.tris[1]
    assert [v(t.p[k]).tolist() for k in range(3)] == [[0, 0, 0], [8, 8, 0], [0, 8, 0]]
    assert d.materials[p0.mtl].diffuse.x == 1.0 and d.materials[p0.mtl].shiness == 20
    assert d.materials[p1.mtl].diffuse.z == 1.0
    # newmtl does not reset the running material: blue keeps Ka and Ns of red
    assert d.materials[p1.mtl].ambient.x == np.float32(0.1) and d.materials[p1.mtl].shiness == 20


def write_bmp(path, w, h, pixels: bytes, off=54, size_image=0):
    """A 24-bit BMP: BITMAPFILEHEADER + BITMAPINFOHEADER, pixel bytes at `off`."""
    import struct
    fh = b"BM" + struct.pack("<IHHI", off + len(pixels), 0, 0, off)
    ih = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, size_image, 2835, 2835, 0, 0)
    path.write_bytes(fh + ih + b"\xee" * (off - 54) + pixels)


def test_bmp_texture_known_texels(tmp_path, monkeypatch):
    """BMP reading (Model.cpp:283-312) through an MTL map_Kd: the texture file
    is the map's base name with extension .bmp, opened from the working
    directory; w x h from the info header; biSizeImage bytes (w h 3 when 0)
    read at bfOffBits, kept as stored (B, G, R, bottom row first).  Texel
    (x, y) is then bytes 3 (y w + x) .. + 2 (3DElement.cpp:443-449)."""
    rng = np.random.default_rng(5)
    w, h = 4, 3
    pix = rng.integers(0, 256, w * h * 3, dtype=np.uint8).tobytes()
    write_bmp(tmp_path / "wood.bmp", w, h, pix, off=70)  # 16 bytes between header and pixels
    w2, h2 = 2, 5
    pix2 = rng.integers(0, 256, w2 * h2 * 3, dtype=np.uint8).tobytes()
    write_bmp(tmp_path / "tile.bmp", w2, h2, pix2, size_image=w2 * h2 * 3)
    (tmp_path / "t.mtl").write_text("newmtl a\nKd 1 1 1\nmap_Kd C:\\assets\\wood.png\n"
                                     "newmtl b\nmap_Kd tile.tga\nnewmtl c\nmap_Kd wood.jpg\n")
    (tmp_path / "t.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\n"
                                     "usemtl a\nf 1/1/1 2/2/1 3/3/1\nusemtl b\nf 1/1/1 2/2/1 3/3/1\n"
                                     "usemtl c\nf 1/1/1 2/2/1 3/3/1\n")
    monkeypatch.chdir(tmp_path)
    hs = rt.HostScene.new()
    i = hs.add_model_obj("t.obj", "t.mtl")
    d = desc(hs)
    o = d.objects[i]
    assert o.part_count == 3
    texs = []
    for k in range(3):
        m = d.materials[d.parts[o.part_begin + k].mtl]
        assert m.tex >= 0
        t = d.textures[m.tex]
        texs.append((t.w, t.h, bytes(np.ctypeslib.as_array(t.bgr, shape=(t.w * t.h * 3,)))))
    assert texs[0] == (w, h, pix) and texs[1] == (w2, h2, pix2)
    assert texs[2] == texs[0]  # wood.jpg -> wood.bmp, already loaded: the same texture
    # a file that is not a bitmap gives no texture, as in the reference
    (tmp_path / "bad.bmp").write_bytes(b"XX" + bytes(60))
    (tmp_path / "bad.mtl").write_text("newmtl a\nmap_Kd bad.bmp\n")
    j = hs.add_model_obj("t.obj", "bad.mtl")
    d = desc(hs)
    assert d.materials[d.parts[d.objects[j].part_begin].mtl].tex == -1


def test_missing_obj_is_an_empty_model(tmp_path):
    hs = rt.HostScene.new()
    i = hs.add_model_obj(str(tmp_path / "none.obj"), str(tmp_path / "none.mtl"))
    assert desc(hs).objects[i].type == 3


def test_scene_descriptor_is_deterministic():
    a = desc(rt.HostScene.workload(1))
    b = desc(rt.HostScene.workload(1))
    for k in range(a.n_objects):
        assert bytes(a.objects[k]) == bytes(b.objects[k])
    assert bytes(a.lights[0]) == bytes(b.lights[0])


def test_png_writer_round_trip(tmp_path):
    """rtg_host_write_png: the frame decodes back (standard zlib, chunk CRCs)
    to the same pixels, with the reference's bottom-up rows flipped."""
    import struct
    import zlib
    rng = np.random.default_rng(3)
    for (h, w) in ((1, 1), (7, 5), (200, 300)):  # 300 x 200: three stored deflate blocks
        img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        for bottom_up in (False, True):
            path = str(tmp_path / f"f_{h}_{w}_{int(bottom_up)}.png")
            rt.write_png(path, img, bottom_up=bottom_up)
            data = open(path, "rb").read()
            assert data[:8] == b"\x89PNG\r\n\x1a\n"
            pos, chunks = 8, []
            while pos < len(data):
                (n,) = struct.unpack(">I", data[pos:pos + 4])
                typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
                (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
                assert zlib.crc32(typ + body) == crc, typ
                chunks.append((typ, body))
                pos += 12 + n
            assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
            assert struct.unpack(">IIBBBBB", chunks[0][1]) == (w, h, 8, 2, 0, 0, 0)
            raw = zlib.decompress(chunks[1][1])
            rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
            assert (rows[:, 0] == 0).all()
            got = rows[:, 1:].reshape(h, w, 3)
            assert (got == (img[::-1] if bottom_up else img)).all()


def test_reused_output_buffers_are_checked():
    """render_batch / DeviceScene.render with out=: buffers of the wrong shape,
    type or count are refused before anything is launched."""
    p = rt.params(64, 32)

    class _NoScene:
        h = None

    good = {"rgb": np.empty((32, 64, 3), np.uint8), "ids": np.empty((32, 64, 2), np.int32)}
    with pytest.raises(ValueError):
        rt.render_batch([(_NoScene(), p)], out=[])
    with pytest.raises(ValueError):
        rt.render_batch([(_NoScene(), p)], ids=False, out=[{"rgb": np.empty((32, 64, 4), np.uint8), "ids": None}])
    with pytest.raises(ValueError):
        rt.render_batch([(_NoScene(), p)], out=[{"rgb": good["rgb"], "ids": np.empty((32, 64, 2), np.int64)}])
    with pytest.raises(ValueError):
        rt.render_batch([(_NoScene(), p)], ids=False, out=[{"rgb": np.empty((32, 128, 3), np.uint8)[:, ::2], "ids": None}])


def _light(hs, k):
    out = _capi.rtg_light()
    _capi.check(_capi.load().rtg_host_light(hs.h, k, C.byref(out)), "rtg_host_light")
    return out


def test_scene_mutators_closed_forms():
    """Scene edits (Scene.cpp:157-301) against what they must produce:
    ChgLightComp keeps every channel's ambient + diffuse + specular at the
    luminance while weighting the parts; LUMI scales all parts; ChgMtl(colour)
    keeps r + g + b (w < 0.5) or replaces the colour; the 'wall' material
    clears a plane's checker; Delete removes exactly one entry; the ninth
    light is refused."""
    L = _capi.load()
    hs = rt.HostScene.new()
    hs.add_light(2, (0.2, 0.5, 0.3), (1, 0, 0, 2.0))
    a = _light(hs, 0)
    for c in "xyz":  # AddLight: the parts sum to the luminance
        s = getattr(a.ambient, c) + getattr(a.diffuse, c) + getattr(a.specular, c)
        assert abs(s - 2.0) < 1e-6
    assert L.rtg_host_chg_light_comp(hs.h, 1, 0, _capi.floats(1.0, 2.0, 0.5, 0.0)) == 0
    b = _light(hs, 0)
    for c in "xyz":
        pa, pd, ps = (getattr(a.ambient, c), getattr(a.diffuse, c) * 2.0, getattr(a.specular, c) * 0.5)
        tot = pa + pd + ps
        want = np.array([pa, pd, ps]) / tot * 2.0
        got = np.array([getattr(b.ambient, c), getattr(b.diffuse, c), getattr(b.specular, c)])
        assert np.allclose(got, want, rtol=1e-6), (c, got, want)
    assert L.rtg_host_chg_light_comp(hs.h, 2, 0, _capi.floats(0, 0, 0, 1.5)) == 0
    c_ = _light(hs, 0)
    assert abs(c_.attenuation.w - 3.0) < 1e-6 and abs(c_.diffuse.y - b.diffuse.y * 1.5) < 1e-6
    assert L.rtg_host_chg_light_comp(hs.h, 3, 0, _capi.floats(1, 1, 1, 1)) != 0  # unknown type
    assert L.rtg_host_chg_light_comp(hs.h, 1, 5, _capi.floats(1, 1, 1, 1)) != 0  # no such light
    for k in range(7):
        assert hs.add_light(2, (1, 1, 1), (1, 0, 0, 1)) == k + 1
    assert hs.add_light(2, (1, 1, 1), (1, 0, 0, 1)) == 255
    # recolour
    s = hs.add_sphere(1.0)  # bas-sphere diffuse (0.1, 0.5, 0.8)
    assert L.rtg_host_recolour(hs.h, s, _capi.floats(1.0, 0.0, 0.0, 0.0)) == 0
    m = desc(hs).materials[desc(hs).objects[s].mtl]
    assert abs(m.diffuse.x - 1.4) < 1e-6 and m.diffuse.y == 0 and m.diffuse.z == 0
    assert L.rtg_host_recolour(hs.h, s, _capi.floats(0.25, 0.5, 0.75, 1.0)) == 0
    m = desc(hs).materials[desc(hs).objects[s].mtl]
    assert (m.diffuse.x, m.diffuse.y, m.diffuse.z) == (0.25, 0.5, 0.75)
    # the wall material clears a plane's checker; other objects keep theirs
    p = hs.add_plane()
    t0 = desc(hs).textures[desc(hs).objects[p].tex]
    assert len(set(np.ctypeslib.as_array(t0.bgr, shape=(48,)).tolist())) == 2
    assert L.rtg_host_chg_mtl_library(hs.h, p, 5) == 0
    t1 = desc(hs).textures[desc(hs).objects[p].tex]
    assert set(np.ctypeslib.as_array(t1.bgr, shape=(48,)).tolist()) == {255}
    # delete
    n = desc(hs).n_objects
    assert L.rtg_host_delete(hs.h, 2, 0) == 0 and desc(hs).n_objects == n - 1
    assert desc(hs).objects[0].type == 4  # the plane moved down
    assert L.rtg_host_delete(hs.h, 2, 9) != 0
    assert L.rtg_host_delete(hs.h, 1, 0) == 0
