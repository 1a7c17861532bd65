// intersect.cpp — DrawObject::intersect of the drop-in classes: one ray
// against one object, on the host (3DElement.h:201 and its overrides in
// Basic3DObject.cpp:135-190, 274-300, 376-410, 486-552 and Model.cpp:748-811),
// plus the octant bins Model::intersect walks (Model::RTPrepare,
// Model.cpp:402-480).
//
// The distance tests are the device kernels' own (../csrc/rt_core.hpp,
// compiled here for the host), so a caller's query and the renderer agree;
// the hit record is assembled with the drop-in Vertex / Normal arithmetic,
// whose rounding follows the reference's SSE code.  The renderer itself never
// calls these: it intersects on the device.
#include "Basic3DObject.h"
#include "Model.h"
#include "../csrc/rt_core.hpp"

namespace {

rtg::V3 v3(const Vertex &v) { return rtg::V3{v.x, v.y, v.z}; }
rtg::Ray dev_ray(const Ray &r) { return rtg::Ray{v3(r.origin), v3(r.direction)}; }
intptr_t id_of(const void *p) { return reinterpret_cast<intptr_t>(p); }

// A primitive's distance replaces the running hit when it is nearer and beyond
// 1e-6 (the double constant: t > 1e-6 is compared in double).
bool accepted(float t, float limit) { return t < limit && t > 1e-6; }

// A record at distance t along `ray` naming `who` and its material.
HitRes record(const Ray &ray, float t, Material *m, intptr_t who) {
    HitRes h(t);
    h.position = ray.origin + ray.direction * t;
    h.mtl = m;
    h.obj = who;
    return h;
}

// Outcome of one sphere (a Sphere, or one of a BallPlane's 16) for a ray
// whose previous hit is `hr`, against the running best distance `limit`.
enum class SphereStep { None, Hit, ShadowInside };

// The sphere the ray leaves (`leaving`): from outside it is skipped; from
// inside, a shadow ray is blocked by it (at the distance of the radius) and
// any other ray meets its far side with the inward normal, IOR 1 for a
// refraction ray stepping out and the material's otherwise.  Any other
// sphere: the near side with the outward normal.
SphereStep sphere_step(const Ray &ray, bool leaving, const Vertex &c, float rsq, Material &m, intptr_t who,
                       float limit, HitRes &out) {
    const rtg::Ray r = dev_ray(ray);
    if (leaving) {
        if (!ray.isInside) return SphereStep::None;
        if (ray.type == MY_RAY_SHADOWRAY) return SphereStep::ShadowInside;
        const float t = rtg::sphere_exit_t(r, v3(c), rsq);
        if (!accepted(t, limit)) return SphereStep::None;
        out = record(ray, t, &m, who);
        out.normal = Normal(c - out.position);
        out.isInside = (uint8_t)~ray.isInside;
        out.rfr = ray.type == MY_RAY_REFRACTRAY ? 1.0f : m.rfr;
        return SphereStep::Hit;
    }
    const float t = rtg::sphere_enter_t(r, v3(c), rsq);
    if (t == rtg::kMiss || !accepted(t, limit)) return SphereStep::None;  // pointing away, or no root
    out = record(ray, t, &m, who);
    out.normal = Normal(out.position - c);
    out.isInside = (uint8_t)~ray.isInside;
    out.rfr = m.rfr;
    return SphereStep::Hit;
}

}  // namespace

HitRes Sphere::intersect(const Ray &ray, const HitRes &hr, const float) {
    HitRes out;
    switch (sphere_step(ray, hr.obj == id_of(this), position, radius_sqr, mtl, id_of(this), hr.distance, out)) {
    case SphereStep::Hit:
        return out;
    case SphereStep::ShadowInside:
        return HitRes(radius);
    default:
        return hr;
    }
}

// The 4 x 4 spheres sit at position + cx axisx + cy axisy for cx, cy stepping
// from -6r by 4r while below 8r, cx outermost; sphere k (1..16, in that order)
// answers to the identity this + k.
HitRes BallPlane::intersect(const Ray &ray, const HitRes &hr, const float) {
    HitRes best = hr;
    int k = 0;
    for (float cx = radius * -6; cx < radius * 8; cx += radius * 4)
        for (float cy = radius * -6; cy < radius * 8; cy += radius * 4) {
            const intptr_t who = id_of(this) + ++k;
            const Vertex c = (axisx * cx + axisy * cy) + position;
            HitRes h;
            const SphereStep s = sphere_step(ray, hr.obj == who, c, radius_sqr, mtl, who, best.distance, h);
            if (s == SphereStep::ShadowInside) return HitRes(radius);
            if (s == SphereStep::Hit) best = h;
        }
    return best.distance < hr.distance ? best : hr;
}

// Infinite plane through `position` with the checker texture, uv from the
// plane frame (u, v = in-plane coordinates / 5); no t epsilon.
HitRes Plane::intersect(const Ray &ray, const HitRes &hr, const float) {
    if (hr.obj == id_of(this)) return hr;
    const rtg::Ray r = dev_ray(ray);
    const float t = rtg::plane_t(r, v3(position), v3(normal));
    if (t == rtg::kMiss || !(t < hr.distance)) return hr;  // parallel, or behind the origin
    HitRes out = record(ray, t, &mtl, id_of(this));
    out.normal = normal;
    out.tex = &tex;
    rtg::plane_uv(r, v3(position), v3(axisx), v3(axisy), out.tcoord.u, out.tcoord.v);
    return out;
}

// Axis-aligned box [min, max] + position; the normal has a component on each
// axis whose coordinate lies within 1e-6 of the half-extent max: an edge or
// corner hit gets a diagonal normal, and a hit that matches no axis (rounding,
// or the inverted box of Box(l, w, h)) gets 0/0.  No inside path: a ray that
// starts inside has entry 0 and misses.
HitRes Box::intersect(const Ray &ray, const HitRes &hr, const float) {
    if (hr.obj == id_of(this)) return hr;
    const float t = rtg::border_test(dev_ray(ray), v3(min + position), v3(max + position));
    if (!accepted(t, hr.distance)) return hr;
    HitRes out = record(ray, t, &mtl, id_of(this));
    const Vertex rel = out.position - position;
    auto side = [](float c, float half) { return std::fabs(std::fabs(c) - half) < 1e-6 ? (c > 0 ? 1.0f : -1.0f) : 0.0f; };
    Vertex face;
    face.z = side(rel.z, max.z);
    face.y = side(rel.y, max.y);
    face.x = side(rel.x, max.x);
    out.normal = Normal(face);
    return out;
}

// Octant bins of every part, in world space: a triangle goes into each bin of
// the part box split at its model-space centre that its own box touches
// (bit 2: upper x, bit 1: upper z, bit 0: upper y); each bin ends with an
// empty sentinel entry.  BorderMin/Max and bboxs are the model and part boxes
// moved to `position`.
void Model::RTPrepare() {
    BorderMin = VerMin + position, BorderMax = VerMax + position;
    bboxs.clear();
    octclparts.assign(parts.size() * 8, {});
    for (size_t a = 0; a < parts.size(); ++a) {
        const Vertex lo = borders[2 * a], hi = borders[2 * a + 1];
        bboxs.push_back(lo + position);
        bboxs.push_back(hi + position);
        const Vertex mid = (lo + hi) * 0.5f;
        for (size_t b = 0; b < parts[a].size(); ++b) {
            const Triangle &t = parts[a][b];
            clTri copy(t.points[1] - t.points[0], t.points[2] - t.points[0], t.points[0] + position);
            copy.numa = (int16_t)a, copy.numb = (int16_t)b;
            const rtg::V3 p0 = v3(t.points[0]), p1 = v3(t.points[1]), p2 = v3(t.points[2]);
            const rtg::V3 mn{rtg::sse_min(p0.x, rtg::sse_min(p1.x, p2.x)), rtg::sse_min(p0.y, rtg::sse_min(p1.y, p2.y)),
                             rtg::sse_min(p0.z, rtg::sse_min(p1.z, p2.z))};
            const rtg::V3 mx{rtg::sse_max(p0.x, rtg::sse_max(p1.x, p2.x)), rtg::sse_max(p0.y, rtg::sse_max(p1.y, p2.y)),
                             rtg::sse_max(p0.z, rtg::sse_max(p1.z, p2.z))};
            const bool half[3][2] = {{mn.x <= mid.x, mx.x >= mid.x},   // x: lower, upper
                                     {mn.z <= mid.z, mx.z >= mid.z},   // z
                                     {mn.y <= mid.y, mx.y >= mid.y}};  // y
            for (int bin = 0; bin < 8; ++bin)
                if (half[0][(bin >> 2) & 1] && half[1][(bin >> 1) & 1] && half[2][bin & 1])
                    octclparts[a * 8 + bin].push_back(copy);
        }
        for (int bin = 0; bin < 8; ++bin) octclparts[a * 8 + bin].push_back(clTri());
    }
}

// Nearest triangle copy over the bins whose half box the ray enters, in part
// and bin order, skipping the copy the ray leaves; the part test compares with
// the incoming hr.distance.  Stops at the first copy closer than `min`.
// Normal and uv are the barycentric blends of the triangle's vertex values.
HitRes Model::intersect(const Ray &ray, const HitRes &hr, const float min) {
    const rtg::Ray r = dev_ray(ray);
    if (!(rtg::border_test(r, v3(BorderMin), v3(BorderMax)) < hr.distance)) return hr;
    float best = hr.distance;
    const clTri *won = nullptr;
    float w = 0, u = 0, v = 0;
    const size_t n_parts = std::min(parts.size(), bboxs.size() / 2);
    for (size_t a = 0; a < n_parts && octclparts.size() >= 8 * (a + 1); ++a) {
        uint32_t mask = 0;
        if (!(rtg::border_test_ex(r, v3(bboxs[2 * a]), v3(bboxs[2 * a + 1]), mask) < hr.distance)) continue;
        for (int bin = 0; bin < 8; ++bin) {
            if (!((mask >> bin) & 1)) continue;
            const vector<clTri> &cell = octclparts[a * 8 + bin];
            for (size_t c = 0; c + 1 < cell.size(); ++c) {
                if (hr.obj == id_of(&cell[c])) continue;
                float cw = 0, cu = 0, cv = 0;
                const float t = rtg::tri_test(r, v3(cell[c].axisu), v3(cell[c].axisv), v3(cell[c].p0), cw, cu, cv);
                if (!(t < best)) continue;
                best = t, won = &cell[c], w = cw, u = cu, v = cv;
                if (t < min) goto done;
            }
        }
    }
done:
    if (won == nullptr) return hr;
    const Triangle &tri = parts[won->numa][won->numb];
    const int8_t m = part_mtl[won->numa];
    HitRes out = record(ray, best, &mtls[m], id_of(won));
    out.normal = tri.norms[0] * w + tri.norms[1] * u + tri.norms[2] * v;
    out.tcoord = tri.tcoords[0] * w + tri.tcoords[1] * u + tri.tcoords[2] * v;
    out.rfr = mtls[m].rfr;
    if (mtl_tex[m] >= 0) out.tex = &texs[mtl_tex[m]];
    return out;
}
