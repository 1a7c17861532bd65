// rt_prep.hpp — host-side flattening of an rtg_scene_desc into the device
// records of rt_scene.hpp.  Shared by the device build (rtg_bvh.hip) and the
// CPU emulation harness of the test suite.
#pragma once
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_scene.hpp"
#include "../../include/rtg.h"

namespace rtg {

struct ScenePrep {
    std::vector<DevObject> objs;
    std::vector<int> brute, spheres;
    std::vector<DevPart> parts;
    std::vector<float4> part_mid;  // model-space border midpoints
    std::vector<int> tri_part;
    std::vector<int> tri_mtl;  // the triangle's part's material
    std::vector<const rtg_triangle *> part_tris;  // per global part, NULL for hidden models
    std::vector<int> part_order;                  // parts in global triangle order
    std::vector<DevMaterial> mtls;
    std::vector<DevLight> lights;
    std::vector<DevTexture> texs;
    std::vector<uint8_t> texels;
    int64_t n_tris = 0;
    V3 env{0, 0, 0};
    DevCamera cam{};
};

inline V3 prep_v3(const rtg_vec4 &a) { return V3{a.x, a.y, a.z}; }

// Returns an empty string on success, else the error message.
inline std::string prepare_scene(const rtg_scene_desc &d, ScenePrep &P) {
    P = ScenePrep();
    P.objs.resize(d.n_objects);
    P.parts.resize(d.n_parts);
    P.part_mid.resize(d.n_parts);
    P.part_tris.assign(d.n_parts, nullptr);
    for (int i = 0; i < d.n_objects; ++i) {
        const rtg_object &o = d.objects[i];
        DevObject &D = P.objs[i];
        memset(&D, 0, sizeof(D));
        D.type = o.type, D.show = o.show, D.mtl = o.mtl, D.tex = o.tex;
        D.pos = prep_v3(o.position);
        D.radius = o.radius, D.rsq = o.radius_sqr;
        D.part_begin = o.part_begin, D.part_count = o.part_count;
        D.normal = prep_v3(o.normal), D.axisx = prep_v3(o.axisx), D.axisy = prep_v3(o.axisy);
        if (o.type == RTG_OBJ_MODEL) {  // BorderMin/BorderMax = VerMin/VerMax + position (Model.cpp:404)
            D.bmin = prep_v3(o.vmin) + D.pos;
            D.bmax = prep_v3(o.vmax) + D.pos;
        } else {
            D.bmin = prep_v3(o.vmin), D.bmax = prep_v3(o.vmax);
        }
        if (!(std::fabs(o.radius) < 1e30f) && (o.type == RTG_OBJ_SPHERE || o.type == RTG_OBJ_BALLPLANE))
            return "object " + std::to_string(i) + " has a non-finite or huge radius";
        if ((o.mtl < 0 || o.mtl >= d.n_materials) && o.type != RTG_OBJ_MODEL)
            return "object " + std::to_string(i) + " has no valid material";
        if (!o.show) continue;
        if (o.type == RTG_OBJ_SPHERE) P.spheres.push_back(i);
        if (o.type == RTG_OBJ_PLANE || o.type == RTG_OBJ_CUBE || o.type == RTG_OBJ_BALLPLANE) P.brute.push_back(i);
        if (o.type != RTG_OBJ_MODEL) continue;
        int first_prim = 0;
        for (int a = 0; a < o.part_count; ++a) {
            const int pi = o.part_begin + a;
            if (pi < 0 || pi >= d.n_parts) return "model part range out of bounds";
            const rtg_part &p = d.parts[pi];
            if (p.tri_count < 0 || (p.tri_count > 0 && p.tris == nullptr)) return "part without triangles";
            if (p.mtl < 0 || p.mtl >= d.n_materials) return "part material out of range";
            DevPart &R = P.parts[pi];
            R.bbmin = prep_v3(p.bmin) + D.pos;  // bboxs = borders + position (Model.cpp:419)
            R.bbmax = prep_v3(p.bmax) + D.pos;
            R.obj = i, R.mtl = p.mtl, R.first_prim = first_prim, R.gtri_base = (int)P.n_tris;
            const V3 mid = (prep_v3(p.bmin) + prep_v3(p.bmax)) * 0.5f;  // Model.cpp:420
            P.part_mid[pi] = make_float4(mid.x, mid.y, mid.z, 0.0f);
            P.part_tris[pi] = p.tris;
            P.part_order.push_back(pi);
            first_prim += p.tri_count;
            P.n_tris += p.tri_count;
            P.tri_part.insert(P.tri_part.end(), (size_t)p.tri_count, pi);
            P.tri_mtl.insert(P.tri_mtl.end(), (size_t)p.tri_count, p.mtl);
        }
    }
    if (P.n_tris + (int64_t)P.spheres.size() >= (1 << 28)) return "more than 2^28 primitives";
    P.mtls.resize(d.n_materials > 0 ? d.n_materials : 1);
    for (int i = 0; i < d.n_materials; ++i) {
        const rtg_material &m = d.materials[i];
        DevMaterial &M = P.mtls[i];
        memset(&M, 0, sizeof(M));
        M.ka = prep_v3(m.ambient), M.kd = prep_v3(m.diffuse), M.ks = prep_v3(m.specular);
        M.shiness = m.shiness, M.reflect = m.reflect, M.refract = m.refract, M.rfr = m.rfr;
        M.tex = m.tex;
    }
    if (d.n_lights > RTG_MAX_LIGHTS) return "more than RTG_MAX_LIGHTS lights";
    P.lights.resize(d.n_lights > 0 ? d.n_lights : 1);
    for (int i = 0; i < d.n_lights; ++i) {
        const rtg_light &l = d.lights[i];
        DevLight &L = P.lights[i];
        L.type = l.type, L.on = l.on;
        L.pos = prep_v3(l.position), L.la = prep_v3(l.ambient), L.ld = prep_v3(l.diffuse);
        L.ls = prep_v3(l.specular), L.att = prep_v3(l.attenuation);
        L.pos_w = l.position.w;
    }
    P.texs.resize(d.n_textures > 0 ? d.n_textures : 1);
    for (int i = 0; i < d.n_textures; ++i) {
        const rtg_texture &t = d.textures[i];
        if (t.w <= 0 || t.h <= 0 || t.bgr == nullptr) return "bad texture";
        P.texs[i] = DevTexture{t.w, t.h, (int)P.texels.size(), 0};
        P.texels.insert(P.texels.end(), t.bgr, t.bgr + (size_t)t.w * t.h * 3);
    }
    if (P.texels.empty()) P.texels.push_back(0);
    P.env = prep_v3(d.env_light);
    P.cam.pos = prep_v3(d.camera.position);
    P.cam.u = prep_v3(d.camera.u);
    P.cam.v = prep_v3(d.camera.v);
    P.cam.n = prep_v3(d.camera.n);
    P.cam.zNear = d.camera.zNear;
    P.cam.zFar = (float)(1.4142135623730951 * (double)d.camera.zFar);  // sqrt(2) * zFar (RayTracer.cpp:15)
    return "";
}

// Level-0 length of a batch of `total` samples, padded so that batches of
// similar size share one captured graph: 4096, then +12.5% steps rounded up
// to a multiple of 4096.  Always a whole number of 64-sample chunks (the
// batch's chunk -> request table has one entry per chunk) and at most 12.5%
// (or 4096 samples) of padding.
inline size_t batch_bucket(size_t total) {
    size_t padded = 4096;
    while (padded < total) padded = (padded + padded / 8 + 4095) & ~size_t(4095);
    return padded;
}

// Most rays one level of a wavefront may hold.  A shadow-ray entry packs the
// index of its hit node with the light index as (node << 5 | light) in 32
// bits, and entries are appended at 32-bit signed offsets: a level holds at
// most 2^27 - 1 rays and at most 2^31 - 1 shadow entries (rays x lights).
inline size_t level_ray_cap(int n_on) {
    const size_t by_node = (size_t(1) << 27) - 1;
    const size_t by_entries = ((size_t(1) << 31) - 1) / (size_t)(n_on > 1 ? n_on : 1);
    return by_node < by_entries ? by_node : by_entries;
}

// Rows of pixel blocks per pass of a single frame (`per_row` level-0
// samples each).  Rays that branch (fanout 2: reflection and refraction) run
// in passes of at most `branch_pass` samples; otherwise level 0 is the
// largest level and a pass takes as many rows as the level cap allows.  Deeper
// levels of a branching frame are checked against the cap as they are
// counted.
inline int frame_rows_per_pass(int rows, size_t per_row, int fanout, int n_on, size_t branch_pass) {
    size_t lim = level_ray_cap(n_on);
    if (fanout >= 2 && branch_pass < lim) lim = branch_pass;
    const size_t r = per_row > 0 ? lim / per_row : (size_t)rows;
    return (int)(r < 1 ? 1 : (r > (size_t)rows ? (size_t)rows : r));
}

// End of the group of consecutive requests that starts at `begin`: the
// longest run whose padded level-0 length (batch_bucket) stays within `cap`.
// A request that alone exceeds the cap is a group of one, rendered as a
// frame in passes.
inline int batch_group_end(const size_t *samples, int n, int begin, size_t cap) {
    size_t acc = samples[begin];
    int i = begin + 1;
    if (batch_bucket(acc) > cap) return i;
    while (i < n && batch_bucket(acc + samples[i]) <= cap) acc += samples[i++];
    return i;
}

}  // namespace rtg
