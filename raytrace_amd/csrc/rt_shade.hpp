// rt_shade.hpp — hit records, local Blinn-Phong shading and the per-sample
// ray-tree evaluation of RTfrac (RayTracer.cpp:450-596).
//
// The colour of a ray-tree node is composed bottom-up in the reference's
// order: c = local; c = c(1-rl) + C_refl rl; c = c(1-rr) + (C_refr . w) rr,
// with w the Beer factor exp(kd 0.15 (-t_child)) when the hit is inside a
// sphere.  Both engines reproduce this order exactly, so their results are
// bit-identical to each other.
#pragma once
#include "rt_trace.hpp"

namespace rtg {

// HitRes fields of a closest hit (Basic3DObject.cpp:151-160,180-187,398-405,
// Model.cpp:796-806).
struct HitRec {
    V3 P, N;
    float tu, tv;
    float t;
    float rfr;
    int mtl, tex;
    int inside;  // HitRes::isInside
    SrcId id;    // identity children skip
};

RTG_HD HitRec make_hitrec(const SceneView &S, const TraceRay &q, const Hit &h) {
    HitRec hr;
    hr.t = h.t;
    hr.P = q.r.o + q.r.d * h.t;
    hr.tu = hr.tv = 0;
    hr.rfr = 1.0f;
    hr.inside = 0;
    hr.tex = -1;
    hr.id = SrcId{h.obj, h.prim, -1};
    const DevObject &o = S.objs[h.obj];
    switch (h.kind) {
    case HIT_TRI: {
        const int g = h.prim;
        const float4 *n = S.tri_nrm + 3 * (size_t)g;
        const float2 *uv = S.tri_uv + 3 * (size_t)g;
        const V3 n0 = V3{n[0].x, n[0].y, n[0].z}, n1 = V3{n[1].x, n[1].y, n[1].z}, n2 = V3{n[2].x, n[2].y, n[2].z};
        hr.N = normalize((n0 * h.w + n1 * h.u) + n2 * h.v);
        hr.tu = (uv[0].x * h.w + uv[1].x * h.u) + uv[2].x * h.v;
        hr.tv = (uv[0].y * h.w + uv[1].y * h.u) + uv[2].y * h.v;
        // the material comes from its own per-triangle table, so that its
        // load does not wait for the part's
        const int mi = S.tri_mtl[g];
        hr.mtl = mi;
        hr.rfr = S.mtls[mi].rfr;
        hr.tex = S.mtls[mi].tex;
        hr.id.bin = record_bin(S, q.r, q.src, g);
        break;
    }
    case HIT_SPHERE: {
        const bool exit = q.src.obj == h.obj;
        hr.N = exit ? normalize(o.pos - hr.P) : normalize(hr.P - o.pos);
        hr.mtl = o.mtl;
        hr.inside = !q.inside;
        hr.rfr = exit && q.type == RAY_REFRACT ? 1.0f : S.mtls[o.mtl].rfr;
        break;
    }
    case HIT_PLANE: {
        hr.N = o.normal;
        hr.mtl = o.mtl;
        hr.tex = o.tex;
        plane_uv(q.r, o.pos, o.axisx, o.axisy, hr.tu, hr.tv);
        break;
    }
    case HIT_CUBE: {
        hr.mtl = o.mtl;
        const V3 b2p = hr.P - o.pos;
        V3 pt = V3{0, 0, 0};
        if (below_1e6(fabsf(b2p.z) - o.bmax.z)) pt.z = b2p.z > 0 ? 1.0f : -1.0f;
        if (below_1e6(fabsf(b2p.y) - o.bmax.y)) pt.y = b2p.y > 0 ? 1.0f : -1.0f;
        if (below_1e6(fabsf(b2p.x) - o.bmax.x)) pt.x = b2p.x > 0 ? 1.0f : -1.0f;
        hr.N = normalize(pt);
        break;
    }
    case HIT_BALL: {
        V3 c;
        ballplane_center(o, h.prim, c);
        const bool exit = q.src.obj == h.obj && q.src.prim == h.prim;
        hr.N = exit ? normalize(c - hr.P) : normalize(hr.P - c);
        hr.mtl = o.mtl;
        hr.inside = !q.inside;
        hr.rfr = exit && q.type == RAY_REFRACT ? 1.0f : S.mtls[o.mtl].rfr;
        break;
    }
    default: break;
    }
    return hr;
}

// Color(const Texture*, const Coord2D&) (3DElement.cpp:430-450); a texel index
// one past the image (coordinate wrapped to exactly 1.0) is clamped.
RTG_HD V3 texel(const SceneView &S, int tex, float u, float v) {
    if (tex < 0 || tex >= S.n_texs) return V3{1.0f, 1.0f, 1.0f};
    const DevTexture &t = S.texs[tex];
    float iu, iv;
    float nu = modff(u, &iu), nv = modff(v, &iv);
    if (nu < 0) nu += 1;
    if (nv < 0) nv += 1;
    const float fx = nu * (float)t.w, fy = nv * (float)t.h;
    const int x = fx == fx ? (int)(int16_t)(int)fx : 0, y = fy == fy ? (int)(int16_t)(int)fy : 0;
    long idx = (long)y * t.w + x;
    const long n = (long)t.w * t.h;
    idx = idx < 0 ? 0 : (idx >= n ? n - 1 : idx);
    const uint8_t *px = S.texels + t.off + 3 * idx;
    return V3{px[2] / 255.0f, px[1] / 255.0f, px[0] / 255.0f};
}

// Per-light quantities of RTfrac's light loop (RayTracer.cpp:482-503).
struct LightTerm {
    V3 la, ld, ls, p2l;
    float dis;
};

RTG_HD LightTerm light_term(const DevLight &L, V3 P) {
    LightTerm lt;
    if (L.type == 2) {  // MY_LIGHT_POINT
        const V3 p2l_v = L.pos - P;
        float dis = dot(p2l_v, p2l_v);
        float step = L.att.x + L.att.z * dis;
        dis = sqrtf(dis);
        step += L.att.y * dis;
        const float lum = 1.0f / step;
        lt.la = L.la * lum;
        lt.ld = L.ld * lum;
        lt.ls = L.ls * lum;
        lt.p2l = normalize(p2l_v);
        lt.dis = dis;
    } else {
        lt.dis = 1e10f;
        lt.la = L.la;
        lt.ld = L.ld;
        lt.ls = L.ls;
        lt.p2l = normalize(L.pos);
    }
    return lt;
}

// Diffuse and specular factors of a light at a hit (RayTracer.cpp:525-540):
// n_l = N.l, n_h = N.h with h the half vector of l and -D.  A light adds
// nothing unless one of them is positive, whether or not it is blocked, so
// its shadow ray need not be traced (it is still counted: the reference
// traces it).
struct LightDots {
    float n_l, n_h;
    RTG_HD bool matters() const { return n_l > 0 || n_h > 0; }
};
RTG_HD LightDots light_dots(V3 N, V3 D, const LightTerm &lt) {
    LightDots d;
    d.n_l = dot(N, lt.p2l);
    d.n_h = dot(N, normalize(lt.p2l - D));
    return d;
}

// Local colour vc (kd...+ka...) + specular (RayTracer.cpp:469-546) at hit
// point P with normal N, texel colour vc, incoming direction D.
// `blocked(li, lt, matters)` answers the shadow test of light li; it is
// called for every switched-on light (so it can count the ray), and its
// answer is used only when `matters`.
template <class Blocked>
RTG_HD V3 shade_local(const SceneView &S, V3 D, V3 P, V3 N, int mtl, V3 vc, Blocked &&blocked) {
    const DevMaterial &m = S.mtls[mtl];
    V3 vd = V3{0, 0, 0}, vs = V3{0, 0, 0};
    V3 va = mixmul(m.ka, S.env);
    for (int li = 0; li < S.n_lights; ++li) {
        const DevLight &L = S.lights[li];
        if (!L.on) continue;
        const LightTerm lt = light_term(L, P);
        va = va + mixmul(m.ka, lt.la);
        const LightDots ld = light_dots(N, D, lt);
        if (blocked(li, lt, ld.matters()) || !ld.matters()) continue;
        if (ld.n_l > 0) vd = vd + mixmul(m.kd, lt.ld) * ld.n_l;
        if (ld.n_h > 0) vs = vs + mixmul(m.ks, lt.ls) * powf(ld.n_h, m.shiness);
    }
    return mixmul(vc, vd + va) + vs;
}

template <class Blocked>
RTG_HD V3 local_color(const SceneView &S, V3 D, const HitRec &hr, Blocked &&blocked) {
    return shade_local(S, D, hr.P, hr.N, hr.mtl, texel(S, hr.tex, hr.tu, hr.tv), blocked);
}


// Child ray of a hit.
RTG_HD TraceRay reflect_ray(const TraceRay &q, const HitRec &hr) {
    const float n_n = 2.0f * dot(q.r.d, hr.N);
    TraceRay c;
    c.r.o = hr.P;
    c.r.d = normalize(q.r.d - hr.N * n_n);
    c.type = RAY_REFLECT;
    c.inside = 0;
    c.src = hr.id;
    c.self_test = 0;
    return c;
}

// Refraction direction; false on total internal reflection (RayTracer.cpp:571-581).
RTG_HD bool refract_ray(const TraceRay &q, float mtlrfr, const HitRec &hr, TraceRay &c) {
    const float n = mtlrfr / hr.rfr;
    const float cosIn = -dot(q.r.d, hr.N);
    const float cosOut2 = 1.0f - (n * n) * (1.0f - cosIn * cosIn);
    if (cosOut2 < 0.0f) return false;
    const V3 l2 = q.r.d * n, l1 = hr.N * (n * cosIn - sqrtf(cosOut2));
    c.r.o = hr.P;
    c.r.d = normalize(l1 + l2);
    c.type = RAY_REFRACT;
    c.inside = hr.inside;
    c.src = hr.id;
    c.self_test = 0;
    return true;
}

// Beer factor of a refracted child (RayTracer.cpp:585-590).
RTG_HD V3 beer(const V3 &kd, float child_alpha) {
    const V3 e = (kd * 0.15f) * (-child_alpha);
    return V3{expf(e.x), expf(e.y), expf(e.z)};
}

// Ray-tree node composition (RayTracer.cpp:546-595) for the wavefront
// engine's bottom-up pass: reflection child c_refl, refraction child c_refr
// with hit distance refr_alpha; `tir` = total internal reflection.  Same
// operations, in the same order, as the megakernel's frame stages.
RTG_HD V3 compose(V3 c, float rl, float rr, bool tir, bool inside, V3 kd, V3 c_refl, V3 c_refr, float refr_alpha) {
    if (rl > 0.01f) {
        c = c * (1.0f - rl);
        c = c + c_refl * rl;
    }
    if (rr > 0.01f) {
        c = c * (1.0f - rr);
        if (!tir) {
            const V3 w = inside ? beer(kd, refr_alpha) : V3{1.0f, 1.0f, 1.0f};
            c = c + mixmul(c_refr, w) * rr;
        }
    }
    return c;
}

struct RayCounts {
    uint32_t primary, shadow, reflect, refract;
    int retests;  // closest-hit queries re-tested in FP64
};

struct PrimaryOut {
    int obj, prim;
    float t;
};

// Reported primitive id of a hit: triangle index within its model (part
// order), BallPlane sphere index, or -1.
RTG_HD int id_prim(const SceneView &S, const Hit &h) {
    if (h.kind == HIT_TRI) {
        const DevPart &P = S.parts[S.tri_part[h.prim]];
        return h.prim - P.gtri_base + P.first_prim;
    }
    return h.kind == HIT_BALL ? h.prim : -1;
}

// One sample of the megakernel: the ray tree of RTfrac evaluated depth first
// with an explicit frame stack; children are composed into their parent in
// the reference's order.
// `no_refract`: the reflect test (RTflec, RayTracer.cpp:332-448), RTfrac
// without its refraction branch.
template <bool kRetest = false, class Stack>
RTG_HD V3 render_sample(const SceneView &S, int max_level, const TraceRay &primary, Stack &st, RayCounts &rc,
                        PrimaryOut &po, bool no_refract = false) {
    struct Frame {
        V3 c, P, N, D;
        float t, bwc, mtlrfr, rfr, rl, rr;
        int mtl, inside, level, stage;
        SrcId id;
    };
    Frame fr[RTG_MAX_DEPTH];
    int sp = 0;
    TraceRay q = primary;
    float q_mtlrfr = 1.0f, q_bwc = 1.0f;
    int q_level = 0;
    V3 ret = V3{0, 0, 0};
    float ret_a = kMiss;
    const float zFar = S.cam.zFar;
    for (;;) {
        // ---- enter the pending ray q
        bool produced = false;
        if (q_level > max_level || q_bwc < 1e-5f) {
            ret = V3{0, 0, 0}, ret_a = kMiss;
        } else {
            if (q_level == 0)
                rc.primary++;
            else if (q.type == RAY_REFLECT)
                rc.reflect++;
            else
                rc.refract++;
            if (q.src.obj >= 0) q.self_test = self_testable(S, q.r, q.src, kMiss);
            const Hit h = trace_closest<kRetest>(S, q, st, &rc.retests);
            if (q_level == 0) po = PrimaryOut{h.obj, id_prim(S, h), h.t};
            const float zNear = q_level == 0 ? S.cam.zNear : 0.0f;
            if (h.kind == HIT_NONE || h.t > zFar || h.t < zNear) {
                ret = V3{0, 0, 0}, ret_a = kMiss;
            } else {
                HitRec hr = make_hitrec(S, q, h);
                const SrcId cs = child_src(S, q, h, st);
                if (cs.obj != -2) hr.id = cs;
                const V3 L = local_color(S, q.r.d, hr, [&](int, const LightTerm &lt, bool matters) -> bool {
                    rc.shadow++;
                    if (!matters) return false;
                    TraceRay sr;
                    sr.r.o = hr.P;
                    sr.r.d = lt.p2l;
                    sr.type = RAY_SHADOW;
                    sr.inside = 0;
                    sr.src = hr.id;
                    sr.self_test = self_testable(S, sr.r, hr.id, lt.dis);
                    return trace_any(S, sr, lt.dis, st);
                });
                Frame &f = fr[sp++];
                f.c = L;
                f.P = hr.P, f.N = hr.N, f.D = q.r.d;
                f.t = hr.t;
                f.bwc = q_bwc;
                f.mtlrfr = q_mtlrfr;
                f.rfr = hr.rfr;
                f.mtl = hr.mtl;
                f.inside = hr.inside;
                f.level = q_level;
                f.stage = 0;
                f.id = hr.id;
                f.rl = S.mtls[hr.mtl].reflect;
                f.rr = no_refract ? 0.0f : S.mtls[hr.mtl].refract;
                produced = true;
            }
        }
        (void)produced;
        // ---- resume parents until one needs a child
        bool descend = false;
        while (sp > 0 && !descend) {
            Frame &f = fr[sp - 1];
            if (f.stage == 1) {  // reflection child returned
                f.c = f.c + ret * f.rl;
                f.stage = 2;
            } else if (f.stage == 3) {  // refraction child returned
                const V3 w = f.inside ? beer(S.mtls[f.mtl].kd, ret_a) : V3{1.0f, 1.0f, 1.0f};
                f.c = f.c + mixmul(ret, w) * f.rr;
                f.stage = 4;
            }
            if (f.stage == 0) {
                f.stage = 2;
                if (f.rl > 0.01f) {
                    f.c = f.c * (1.0f - f.rl);
                    TraceRay qq;
                    qq.r.o = f.P;
                    qq.r.d = normalize(f.D - f.N * (2.0f * dot(f.D, f.N)));
                    qq.type = RAY_REFLECT;
                    qq.inside = 0;
                    qq.src = f.id;
                    qq.self_test = 0;
                    q = qq;
                    q_mtlrfr = 1.0f;
                    q_bwc = f.bwc * f.rl;
                    q_level = f.level + 1;
                    f.stage = 1;
                    descend = true;
                    continue;
                }
            }
            if (f.stage == 2) {
                f.stage = 4;
                if (f.rr > 0.01f) {
                    f.c = f.c * (1.0f - f.rr);
                    HitRec hr;
                    hr.P = f.P, hr.N = f.N, hr.rfr = f.rfr, hr.inside = f.inside, hr.id = f.id;
                    TraceRay in;
                    in.r.o = f.P;  // origin unused by refract_ray
                    in.r.d = f.D;
                    TraceRay qq;
                    if (refract_ray(in, f.mtlrfr, hr, qq)) {
                        q = qq;
                        q_mtlrfr = f.rfr;
                        q_bwc = f.bwc * f.rr;
                        q_level = f.level + 1;
                        f.stage = 3;
                        descend = true;
                        continue;
                    }
                }
            }
            if (f.stage == 4) {
                ret = f.c;
                ret_a = f.t;
                --sp;
            }
        }
        if (!descend) return ret;
    }
}

// Debug render modes (RayTracer.cpp:46-330): one primary closest hit,
// coloured by the mode.  RTdepth: Color::set (3DElement.cpp:451-462), log
// depth between zNear and zFar; RTnorm: 0.5 (N + 1); RTtex: the texel, or
// 0.588 grey without texture; RTmtl: lights without shadows, the light kind
// by position.w > 1e-6 and attenuation x + y sqrt(d) + z d; RTshd: lights
// with shadow rays.  RTmtl and RTshd add the environment ambient after the
// lights (RTfrac adds it first).
template <class Stack>
RTG_HD V3 debug_sample(const SceneView &S, int mode, const TraceRay &q, Stack &st, RayCounts &rc, PrimaryOut &po) {
    rc.primary++;
    const Hit h = trace_closest(S, q, st);
    po = PrimaryOut{h.obj, id_prim(S, h), h.t};
    const float zNear = S.cam.zNear, zFar = S.cam.zFar;
    if (mode == 0x2) {  // MY_MODEL_DEPTHTEST
        const float d = h.t;
        if (d <= zNear) return V3{1.0f, 0.0f, 0.0f};
        if (d >= zFar) return V3{0.0f, 0.0f, 0.0f};
        const float after = logf(d), mx = logf(zFar);
        const float g = (mx - after) / mx;
        return V3{g, g, g};
    }
    if (h.kind == HIT_NONE || h.t > zFar) return V3{0.0f, 0.0f, 0.0f};
    if (h.t < zNear) return V3{1.0f, 1.0f, 1.0f};
    const HitRec hr = make_hitrec(S, q, h);
    if (mode == 0x3)  // MY_MODEL_NORMALTEST
        return V3{0.5f * (hr.N.x + 1.0f), 0.5f * (hr.N.y + 1.0f), 0.5f * (hr.N.z + 1.0f)};
    if (mode == 0x4)  // MY_MODEL_TEXTURETEST
        return hr.tex >= 0 && hr.tex < S.n_texs ? texel(S, hr.tex, hr.tu, hr.tv) : V3{0.588f, 0.588f, 0.588f};
    const DevMaterial &m = S.mtls[hr.mtl];
    const V3 vc = texel(S, hr.tex, hr.tu, hr.tv);
    V3 vd = V3{0, 0, 0}, va = V3{0, 0, 0}, vs = V3{0, 0, 0};
    for (int li = 0; li < S.n_lights; ++li) {
        const DevLight &L = S.lights[li];
        if (!L.on) continue;
        LightTerm lt;
        if (mode == 0x5) {  // MY_MODEL_MATERIALTEST
            if ((double)L.pos_w > 1e-6) {
                const V3 p2l_v = L.pos - hr.P;
                const float d2 = dot(p2l_v, p2l_v);
                const float step = (L.att.x + L.att.y * sqrtf(d2)) + L.att.z * d2;
                const float lum = 1.0f / step;
                lt.la = L.la * lum, lt.ld = L.ld * lum, lt.ls = L.ls * lum;
                lt.p2l = normalize(p2l_v);
            } else {
                lt.la = L.la, lt.ld = L.ld, lt.ls = L.ls;
                lt.p2l = normalize(L.pos);
            }
            lt.dis = 1e10f;
        } else {
            lt = light_term(L, hr.P);
        }
        va = va + mixmul(m.ka, lt.la);
        const LightDots ld = light_dots(hr.N, q.r.d, lt);
        if (mode == 0x6) {  // MY_MODEL_SHADOWTEST
            rc.shadow++;
            if (!ld.matters()) continue;
            TraceRay sr;
            sr.r.o = hr.P;
            sr.r.d = lt.p2l;
            sr.type = RAY_SHADOW;
            sr.inside = 0;
            sr.src = hr.id;
            sr.self_test = self_testable(S, sr.r, hr.id, lt.dis);
            if (trace_any(S, sr, lt.dis, st)) continue;
        }
        if (ld.n_l > 0) vd = vd + mixmul(m.kd, lt.ld) * ld.n_l;
        if (ld.n_h > 0) vs = vs + mixmul(m.ks, lt.ls) * powf(ld.n_h, m.shiness);
    }
    va = va + mixmul(m.ka, S.env);
    return mixmul(vc, vd + va) + vs;
}

}  // namespace rtg
