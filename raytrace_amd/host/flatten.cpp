// flatten.cpp — Scene (reference object graph) -> rtg_scene_desc.
//
// Objects keep Scene::Objects order, which is the closest-hit tie order of the
// reference (RayTracer.cpp:458-465).  Each primitive's Material is copied into
// one global table; a Model contributes all of its `mtls`, and its parts refer
// to them through part_mtl (Model.cpp:800-805).  A plane's texture is attached
// to the plane's material copy (HitRes::tex, Basic3DObject.cpp:401).
#include "flatten.h"

static_assert(sizeof(Triangle) == sizeof(rtg_triangle), "Triangle must match rtg_triangle");
static_assert(sizeof(Vertex) == sizeof(rtg_vec4), "Vertex must match rtg_vec4");

static rtg_vec4 v4(const Vertex &v) { return rtg_vec4{v.x, v.y, v.z, v.w}; }

static rtg_material mat(const Material &m, int32_t tex) {
    rtg_material r{};
    r.ambient = v4(m.ambient);
    r.diffuse = v4(m.diffuse);
    r.specular = v4(m.specular);
    r.emission = v4(m.emission);
    r.shiness = m.shiness;
    r.reflect = m.reflect;
    r.refract = m.refract;
    r.rfr = m.rfr;
    r.tex = tex;
    return r;
}

void SceneFlattener::flatten(const Scene &scene, SceneFlat &out) {
    out.objects.clear();
    out.parts.clear();
    out.materials.clear();
    out.textures.clear();
    out.lights.clear();
    int64_t n_tris = 0;
    for (const DrawObject *obj : scene.Objects) {
        rtg_object o{};
        o.type = obj->type;
        o.show = obj->bShow ? 1 : 0;
        o.position = v4(obj->position);
        o.mtl = -1;
        o.tex = -1;
        switch (obj->type) {
        case MY_OBJECT_SPHERE: {
            const Sphere *s = static_cast<const Sphere *>(obj);
            o.radius = s->radius;
            o.radius_sqr = s->radius_sqr;
            o.mtl = (int32_t)out.materials.size();
            out.materials.push_back(mat(s->mtl, -1));
            break;
        }
        case MY_OBJECT_PLANE: {
            const Plane *p = static_cast<const Plane *>(obj);
            o.normal = v4(p->normal);
            o.axisx = v4(p->axisx);
            o.axisy = v4(p->axisy);
            o.tex = (int32_t)out.textures.size();
            out.textures.push_back(rtg_texture{p->tex.w, p->tex.h, p->tex.data});
            o.mtl = (int32_t)out.materials.size();
            out.materials.push_back(mat(p->mtl, o.tex));
            break;
        }
        case MY_OBJECT_CUBE: {
            const Box *b = static_cast<const Box *>(obj);
            o.vmin = v4(b->min);
            o.vmax = v4(b->max);
            o.mtl = (int32_t)out.materials.size();
            out.materials.push_back(mat(b->mtl, -1));
            break;
        }
        case MY_OBJECT_BALLPLANE: {
            const BallPlane *bp = static_cast<const BallPlane *>(obj);
            o.radius = bp->radius;
            o.radius_sqr = bp->radius_sqr;
            o.normal = v4(bp->normal);
            o.axisx = v4(bp->axisx);
            o.axisy = v4(bp->axisy);
            o.mtl = (int32_t)out.materials.size();
            out.materials.push_back(mat(bp->mtl, -1));
            break;
        }
        case MY_OBJECT_MODEL: {
            const Model *m = static_cast<const Model *>(obj);
            o.vmin = v4(m->VerMin);
            o.vmax = v4(m->VerMax);
            const int32_t tex_base = (int32_t)out.textures.size();
            for (const Texture &t : m->texs) out.textures.push_back(rtg_texture{t.w, t.h, t.data});
            const int32_t mtl_base = (int32_t)out.materials.size();
            for (size_t k = 0; k < m->mtls.size(); ++k) {
                int32_t t = k < m->mtl_tex.size() ? m->mtl_tex[k] : -1;
                out.materials.push_back(mat(m->mtls[k], t >= 0 && (size_t)t < m->texs.size() ? tex_base + t : -1));
            }
            if (m->mtls.empty()) out.materials.push_back(mat(Material(), -1));
            o.part_begin = (int32_t)out.parts.size();
            for (size_t p = 0; p < m->parts.size(); ++p) {
                rtg_part rp{};
                rp.tris = reinterpret_cast<const rtg_triangle *>(m->parts[p].data());
                rp.tri_count = (int32_t)m->parts[p].size();
                int32_t pm = p < m->part_mtl.size() ? m->part_mtl[p] : 0;
                if (pm < 0 || (size_t)pm >= std::max<size_t>(m->mtls.size(), 1)) pm = 0;
                rp.mtl = mtl_base + pm;
                if (2 * p + 1 < m->borders.size()) {
                    rp.bmin = v4(m->borders[2 * p]);
                    rp.bmax = v4(m->borders[2 * p + 1]);
                }
                n_tris += rp.tri_count;
                out.parts.push_back(rp);
            }
            o.part_count = (int32_t)out.parts.size() - o.part_begin;
            o.mtl = mtl_base;
            break;
        }
        default:
            break;
        }
        out.objects.push_back(o);
    }
    for (const Light &l : scene.Lights) {
        rtg_light r{};
        r.type = l.type;
        r.on = l.bLight ? 1 : 0;
        r.position = v4(l.position);
        r.ambient = v4(l.ambient);
        r.diffuse = v4(l.diffuse);
        r.specular = v4(l.specular);
        r.attenuation = v4(l.attenuation);
        out.lights.push_back(r);
    }
    rtg_scene_desc &d = out.desc;
    d = rtg_scene_desc{};
    d.objects = out.objects.data();
    d.n_objects = (int32_t)out.objects.size();
    d.parts = out.parts.data();
    d.n_parts = (int32_t)out.parts.size();
    d.n_tris = n_tris;
    d.materials = out.materials.data();
    d.n_materials = (int32_t)out.materials.size();
    d.textures = out.textures.data();
    d.n_textures = (int32_t)out.textures.size();
    d.lights = out.lights.data();
    d.n_lights = (int32_t)out.lights.size();
    d.env_light = v4(scene.EnvLight);
    const Camera &c = scene.cam;
    d.camera.position = v4(c.position);
    d.camera.u = v4(c.u);
    d.camera.v = v4(c.v);
    d.camera.n = v4(c.n);
    d.camera.width = c.width;
    d.camera.height = c.height;
    d.camera.fovy = c.fovy;
    d.camera.zNear = c.zNear;
    d.camera.zFar = c.zFar;
}

namespace {
struct Hash4 {
    uint64_t h[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull, 0x27D4EB2F165667C5ull};
    static uint64_t mix(uint64_t a, uint64_t b) {
        a ^= b + 0x9E3779B97F4A7C15ull + (a << 6) + (a >> 2);
        a *= 0xff51afd7ed558ccdull;
        return a ^ (a >> 33);
    }
    void bytes(const void *p, size_t n) {
        const uint8_t *c = static_cast<const uint8_t *>(p);
        size_t i = 0;
        for (; i + 32 <= n; i += 32)
            for (int k = 0; k < 4; ++k) {
                uint64_t w;
                memcpy(&w, c + i + 8 * k, 8);
                h[k] = (h[k] ^ w) * 0x100000001B3ull;
                h[k] ^= h[k] >> 29;
            }
        uint64_t tail = n;
        for (; i < n; ++i) tail = tail * 131 + c[i];
        h[0] = mix(h[0], tail);
    }
    uint64_t done() const { return mix(mix(h[0], h[1]), mix(h[2], h[3])); }
};
}  // namespace

uint64_t rtg_desc_fingerprint(const rtg_scene_desc &d) {
    Hash4 h;
    h.bytes(d.objects, sizeof(rtg_object) * d.n_objects);
    for (int i = 0; i < d.n_parts; ++i) {
        const rtg_part &p = d.parts[i];
        h.bytes(&p.tri_count, sizeof(int32_t) * 2);
        h.bytes(&p.bmin, sizeof(rtg_vec4) * 2);
        h.bytes(p.tris, sizeof(rtg_triangle) * p.tri_count);
    }
    h.bytes(d.materials, sizeof(rtg_material) * d.n_materials);
    for (int i = 0; i < d.n_textures; ++i) {
        h.bytes(&d.textures[i].w, 8);
        h.bytes(d.textures[i].bgr, (size_t)d.textures[i].w * d.textures[i].h * 3);
    }
    h.bytes(d.lights, sizeof(rtg_light) * d.n_lights);
    h.bytes(&d.env_light, sizeof(rtg_vec4));
    h.bytes(&d.camera, sizeof(rtg_camera));
    return h.done();
}
