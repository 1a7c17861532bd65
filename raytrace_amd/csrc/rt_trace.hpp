// rt_trace.hpp — closest-hit and any-hit queries over a SceneView.
//
// Semantics are those of the reference's per-object scan (RayTracer.cpp:
// 458-465 and 510-520): the closest hit is the smallest t, ties going to the
// lower object index (an object replaces the running hit only when strictly
// closer), and each primitive keeps its own acceptance rule and self-skip
// rule.  The BVH only culls; its boxes are conservative (FP64-expanded leaf
// bounds, child boxes quantised outward, robust slab test), so the result
// equals the scan's.
//
// Self-skip of triangles.  The reference tests the clTri *copies* of a
// model's octant bins and skips only the copy whose address is the ray's
// source (Model.cpp:775); a triangle stored in several bins can therefore be
// hit again by its own secondary rays.  `self_testable` reproduces that: the
// source triangle is tested iff the ray's BorderTest / BorderTestEx reach a
// bin holding another copy of it.
#pragma once
#include "rt_scene.hpp"

namespace rtg {

enum : int { RAY_BASE = 1, RAY_SHADOW = 2, RAY_REFLECT = 3, RAY_REFRACT = 4 };
enum : int { OBJ_SPHERE = 1, OBJ_CUBE = 2, OBJ_MODEL = 3, OBJ_PLANE = 4, OBJ_BALLPLANE = 5 };
enum : int { HIT_NONE = 0, HIT_TRI = 1, HIT_SPHERE = 2, HIT_PLANE = 3, HIT_CUBE = 4, HIT_BALL = 5 };

struct TraceRay {
    Ray r;
    int type;
    int inside;     // Ray::isInside
    SrcId src;      // HitRes::obj of the parent hit
    int self_test;  // source triangle may be hit again (see header)
};

struct Hit {
    float t;
    int kind;
    int obj, prim;
    float w, u, v;  // barycentrics (triangles)
};

RTG_HD uint64_t hit_key(int obj, int prim) { return ((uint64_t)(uint32_t)obj << 32) | (uint32_t)(prim + 1); }

// 1 + 2*gamma(3) with margin (Ize 2013): scaling the far slab distance makes
// the FP32 slab test conservative.
constexpr float kFarScale = 1.0000006f;

RTG_HD V3 safe_inv(V3 d) {
    return V3{d.x != 0.0f ? 1.0f / d.x : copysignf(1e30f, d.x), d.y != 0.0f ? 1.0f / d.y : copysignf(1e30f, d.y),
              d.z != 0.0f ? 1.0f / d.z : copysignf(1e30f, d.z)};
}

RTG_HD bool slab(V3 o, V3 inv, float lx, float hx, float ly, float hy, float lz, float hz, float tmax,
                 float &tnear) {
    float ax = (lx - o.x) * inv.x, bx = (hx - o.x) * inv.x;
    float ay = (ly - o.y) * inv.y, by = (hy - o.y) * inv.y;
    float az = (lz - o.z) * inv.z, bz = (hz - o.z) * inv.z;
    float n = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
    float f = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)) * kFarScale;
    tnear = n;
    return n <= fminf(f, tmax);
}

// The slab test of a ray against the children of one 4-wide node, at most as
// strict as `slab` on the decoded boxes (it never culls a child `slab` keeps).
// Each axis's entry and exit bound is picked by the direction's sign, and its
// distance is one fused multiply-add, q (s inv) + (org - o) inv, with s inv
// exact (s is a power of two).  The decoded bound org + q s = (m + q) s is a
// float exactly (quantize_axis), so against `slab`'s fl(fl(org + q s - o) inv)
// the result is off by at most 6.1 eps |(org - o) inv| + 766 eps |s inv|
// (eps = 2^-24, the constant term's own rounding included); the constant term
// moves out by 16 eps |(org - o) inv| + 2048 eps |s inv|, and rounding to
// nearest keeps the margin's effect monotone.  An overflow makes a term
// infinite or NaN, which fmaxf / fminf drop: the test then only passes more
// children.
struct Node4Test {
    float sx, sy, sz;  // s inv per axis
    float nx, fx, ny, fy, nz, fz;  // (org - o) inv moved out, for entry and exit
    uint32_t qnx, qfx, qny, qfy, qnz, qfz;  // byte words of the entry and exit bounds
    RTG_HD Node4Test(const Bvh4Node &n, V3 o, V3 inv) {
        const uint32_t e = __builtin_bit_cast(uint32_t, n.org.w);
        sx = pow2f(e & 0xFF) * inv.x, sy = pow2f((e >> 8) & 0xFF) * inv.y, sz = pow2f((e >> 16) & 0xFF) * inv.z;
        const float bx = (n.org.x - o.x) * inv.x, by = (n.org.y - o.y) * inv.y, bz = (n.org.z - o.z) * inv.z;
        const float mx = fabsf(bx) * 9.5367431640625e-07f + fabsf(sx) * 1.220703125e-04f;  // 2^-20, 2^-13
        const float my = fabsf(by) * 9.5367431640625e-07f + fabsf(sy) * 1.220703125e-04f;
        const float mz = fabsf(bz) * 9.5367431640625e-07f + fabsf(sz) * 1.220703125e-04f;
        nx = bx - mx, fx = bx + mx, ny = by - my, fy = by + my, nz = bz - mz, fz = bz + mz;
        // a negative direction enters a box through its upper bound
        const bool neg_x = inv.x < 0.0f, neg_y = inv.y < 0.0f, neg_z = inv.z < 0.0f;
        qnx = neg_x ? n.qxy.y : n.qxy.x, qfx = neg_x ? n.qxy.x : n.qxy.y;
        qny = neg_y ? n.qxy.w : n.qxy.z, qfy = neg_y ? n.qxy.z : n.qxy.w;
        qnz = neg_z ? n.qz.y : n.qz.x, qfz = neg_z ? n.qz.x : n.qz.y;
    }
    // Child k reaches the segment [0, tmax]; `near` gets its entry distance.
    RTG_HD bool child(int k, float tmax, float &near) const {
        const int sh = 8 * k;
        const float tnx = fmaf((float)((qnx >> sh) & 0xFF), sx, nx);
        const float tfx = fmaf((float)((qfx >> sh) & 0xFF), sx, fx);
        const float tny = fmaf((float)((qny >> sh) & 0xFF), sy, ny);
        const float tfy = fmaf((float)((qfy >> sh) & 0xFF), sy, fy);
        const float tnz = fmaf((float)((qnz >> sh) & 0xFF), sz, nz);
        const float tfz = fmaf((float)((qfz >> sh) & 0xFF), sz, fz);
        near = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
        const float far = fminf(fminf(tfx, tfy), tfz) * kFarScale;
        return near <= fminf(far, tmax);
    }
};

// Host emulation stack (the device uses an LDS-backed stack, rtg_kernels.hip).
struct ArrayStack {
    int d[128];
    int sp = 0;
    RTG_HD void reset() { sp = 0; }
    RTG_HD void push(int v) { d[sp++] = v; }
    RTG_HD int pop() { return d[--sp]; }
    RTG_HD bool empty() const { return sp == 0; }
    RTG_HD void on_node() {}
    RTG_HD void on_prim() {}
    template <int N, int D>
    RTG_HD bool lanes_at_least(bool v) const {  // one lane
        return v;
    }
    RTG_HD Bvh4Node node(const SceneView &S, int i) const { return S.nodes[i]; }
};

constexpr int kWalkDone = (int)0x80000000;  // no node left: below every leaf code

// Visit the primitives of every leaf the segment [0, tmax] reaches, nearest
// child first (kOrdered; otherwise in child order, for the any-hit test).
// visit(prim_index, tmax&) returns true to stop.  The stack is
// emptied on entry: an any-hit walk that stops early leaves entries behind,
// and the stack only holds one entry per level of the current path (a tree
// of depth < kStackMax) when every walk starts from an empty stack.
//
// kSpec: speculative traversal (Aila and Laine 2009).  A lane that reaches a
// leaf postpones it and keeps walking nodes, postponing a second leaf too,
// until it reaches a third; the node loop ends when enough lanes of the wave
// hold a leaf (Stack::lanes_at_least: half of them in the closest-hit walk,
// three quarters in the any-hit walk), then the held leaves are tested.  The
// lanes of a wave then spend fewer steps waiting for one another.  A single
// lane (lanes_at_least(v) == v) visits in exactly the order of the plain
// walk.  Only for walks whose result does not depend on the order leaves are
// tested in (the any-hit test; the closest hit, whose winner is the least
// (t, key)).
template <bool kSpec = false, bool kOrdered = true, class Stack, class Visit>
RTG_HD void bvh_walk(const SceneView &S, const Ray &r, float &tmax, Stack &st, Visit &&visit) {
    st.reset();
    if (S.n_prims == 0) return;
    const V3 inv = safe_inv(r.d);
    int node = S.root;
    // One 4-wide node: the next node to enter (the nearest child the segment
    // reaches), the other children it reaches pushed farthest first; the
    // popped entry when it reaches none.
    auto step = [&](int nd) -> int {
        st.on_node();
        const Bvh4Node n = st.node(S, nd);
        const int nc = node4_count(n);
        const int lk[4] = {n.link.x, n.link.y, n.link.z, n.link.w};
        float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool h[4];
        const Node4Test nt(n, r.o, inv);
        // a node has 2 to 4 children; all four are tested (the bytes of a
        // missing child decode to a point) and the count masks the result,
        // which keeps the wave's four tests free of branches
#pragma unroll
        for (int k = 0; k < 4; ++k) h[k] = nt.child(k, tmax, t[k]) & (k < nc);
        int next = kWalkDone;
        if constexpr (kOrdered) {
            // Sort keys: the entry distance's bits (t >= 0, so they order as
            // the distances do) above the child's link, or ~0 for a child the
            // segment misses; five compare-exchanges sort them, nearest first.
            // The links travel in the keys: picking them by index afterwards
            // would send the node to scratch memory.  Only the order of the
            // walk depends on the sort.
            uint64_t s[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                s[k] = h[k] ? ((uint64_t)__builtin_bit_cast(uint32_t, t[k] > 0.0f ? t[k] : 0.0f) << 32 |
                               (uint32_t)lk[k])
                            : ~0ull;
            auto cx = [](uint64_t &a, uint64_t &b) {
                const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
                a = lo, b = hi;
            };
            cx(s[0], s[1]), cx(s[2], s[3]), cx(s[0], s[2]), cx(s[1], s[3]), cx(s[1], s[2]);
#pragma unroll
            for (int r = 3; r >= 1; --r)
                if (s[r] != ~0ull) st.push((int)(uint32_t)s[r]);
            if (s[0] != ~0ull) next = (int)(uint32_t)s[0];
        } else {
#pragma unroll
            for (int k = 3; k >= 0; --k)
                if (h[k]) {
                    if (next != kWalkDone) st.push(next);
                    next = lk[k];
                }
        }
        if (next != kWalkDone) return next;
        return st.empty() ? kWalkDone : st.pop();
    };
    if constexpr (!kSpec) {
        for (;;) {
            while (node >= 0) {
                node = step(node);
                if (node == kWalkDone) return;
            }
            const int first = leaf_first(node), count = leaf_count(node);
            for (int i = 0; i < count; ++i) {
                st.on_prim();
                if (visit(first + i, tmax)) return;
            }
            if (st.empty()) return;
            node = st.pop();
        }
    } else {
        int leaf = 0, leaf2 = 0;  // postponed leaves (< 0), or 0
        for (;;) {
            while (node >= 0) {
                node = step(node);
                if (node < 0 && node != kWalkDone && leaf2 == 0) {
                    (leaf == 0 ? leaf : leaf2) = node;
                    node = st.empty() ? kWalkDone : st.pop();
                }
                // closest hit: half the lanes; any hit: three quarters
                if (kOrdered ? st.template lanes_at_least<1, 2>(leaf != 0)
                             : st.template lanes_at_least<3, 4>(leaf != 0))
                    break;
            }
            if (leaf == 0 && node < 0 && node != kWalkDone) {
                leaf = node;
                node = st.empty() ? kWalkDone : st.pop();
            }
            while (leaf < 0) {
                const int first = leaf_first(leaf), count = leaf_count(leaf);
                for (int i = 0; i < count; ++i) {
                    st.on_prim();
                    if (visit(first + i, tmax)) return;
                }
                leaf = leaf2;
                leaf2 = 0;
                if (leaf == 0 && node < 0 && node != kWalkDone) {
                    leaf = node;
                    node = st.empty() ? kWalkDone : st.pop();
                }
            }
            if (node == kWalkDone) return;
        }
    }
}

// Copy bin the reference records for a hit of ray `r` (source `src`) on
// triangle g: the first bin in scan order that the ray's BorderTestEx mask
// reaches and that holds a copy not skipped by the ray (Model.cpp:765-789).
RTG_HD int record_bin(const SceneView &S, const Ray &r, const SrcId &src, int g) {
    const DevPart &P = S.parts[S.tri_part[g]];
    uint32_t K;
    border_test_ex(r, P.bbmin, P.bbmax, K);
    const uint32_t member = S.tri_bins[g];
    uint32_t m = member & K;
    if (src.prim == g && src.obj == P.obj) m &= ~(1u << src.bin);
    if (m == 0) m = member;
    return m ? __builtin_ctz(m) : 0;
}

// Whether a secondary ray may hit its own source triangle (see header).
RTG_HD int self_testable(const SceneView &S, const Ray &r, const SrcId &src, float dist_in) {
    if (src.obj < 0 || src.prim < 0 || S.objs[src.obj].type != OBJ_MODEL) return 0;
    const DevObject &O = S.objs[src.obj];
    if (!(border_test(r, O.bmin, O.bmax) < dist_in)) return 0;
    const DevPart &P = S.parts[S.tri_part[src.prim]];
    uint32_t K;
    if (!(border_test_ex(r, P.bbmin, P.bbmax, K) < dist_in)) return 0;
    return (S.tri_bins[src.prim] & K & ~(1u << src.bin)) != 0;
}

// Centre of BallPlane sphere `k` (0..15), by the reference's float loop
// (Basic3DObject.cpp:492-497).
RTG_HD bool ballplane_center(const DevObject &o, int k, V3 &c) {
    int cnt = 0, ix = 0;
    for (float cx = o.radius * -6; cx < o.radius * 8 && ix < 4; cx += o.radius * 4, ++ix) {
        int iy = 0;
        for (float cy = o.radius * -6; cy < o.radius * 8 && iy < 4; cy += o.radius * 4, ++iy) {
            if (cnt == k) {
                c = (o.axisx * cx + o.axisy * cy) + o.pos;
                return true;
            }
            ++cnt;
        }
    }
    return false;
}

// Objects outside the BVH, in object order.  Returns true for an any-hit
// query when blocked.
RTG_HD bool brute_objects(const SceneView &S, const TraceRay &q, bool anyhit, float dis, Hit &best,
                          int before = 0x7fffffff) {
    for (int k = 0; k < S.n_brute; ++k) {
        const int oi = S.brute[k];
        if (oi >= before) break;  // `brute` is in object order
        const DevObject &o = S.objs[oi];
        const float lim = anyhit ? dis : best.t;
        if (o.type == OBJ_PLANE) {
            if (q.src.obj == oi) continue;
            float t = plane_t(q.r, o.pos, o.normal);
            if (t < lim) {
                if (anyhit) return true;
                best = Hit{t, HIT_PLANE, oi, -1, 0, 0, 0};
            }
        } else if (o.type == OBJ_CUBE) {
            if (q.src.obj == oi) continue;
            float t = border_test(q.r, o.bmin + o.pos, o.bmax + o.pos);
            if (t < lim && t > 1e-6f) {
                if (anyhit) return true;
                best = Hit{t, HIT_CUBE, oi, -1, 0, 0, 0};
            }
        } else if (o.type == OBJ_BALLPLANE) {
            int cnt = 0, ix = 0;
            float run = lim;
            // 4 x 4 spheres; the iteration caps only matter for radii whose
            // float loop would not terminate (|r| near FLT_MAX).
            for (float cx = o.radius * -6; cx < o.radius * 8 && ix < 4; cx += o.radius * 4, ++ix) {
                int iy = 0;
                for (float cy = o.radius * -6; cy < o.radius * 8 && iy < 4; cy += o.radius * 4, ++iy) {
                    const int k2 = cnt++;
                    const V3 c = (o.axisx * cx + o.axisy * cy) + o.pos;
                    float t;
                    if (q.src.obj == oi && q.src.prim == k2) {
                        if (!q.inside) continue;
                        if (anyhit) {
                            if (o.radius < dis) return true;
                            continue;
                        }
                        t = sphere_exit_t(q.r, c, o.rsq);
                    } else {
                        t = sphere_enter_t(q.r, c, o.rsq);
                    }
                    if (t < run && t > 1e-6f) {
                        if (anyhit) return true;
                        run = t;
                        best = Hit{t, HIT_BALL, oi, k2, 0, 0, 0};
                    }
                }
            }
        }
    }
    return false;
}

// ---- FP64 near-tie re-test (RTG_FLAG_FP64_RETEST) --------------------------
//
// A closest-hit query can be decided by FP32 rounding: two candidates within
// 4 ulp of each other, a barycentric or t within 1e-6 of its accept threshold,
// a ray tangent to a sphere.  With the re-test, the walk collects such
// candidates (the same band the oracle flags) and, when there are any,
// re-tests them and the FP32 winner in FP64 from the exact model-space
// vertices (object position added in double); the FP64 closest wins, ties to
// the lower (object, primitive).  Hits outside the band keep their FP32
// result, so frames differ from the reference algorithm only inside the band.

struct D3 {
    double x, y, z;
};
RTG_HD D3 d3(V3 a) { return D3{a.x, a.y, a.z}; }
RTG_HD D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RTG_HD D3 dadd(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RTG_HD double ddot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + 0.0); }
RTG_HD D3 dcross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// tri_test in double (same steps and accept rules).
RTG_HD double tri_test64(D3 o, D3 d, D3 p0, D3 e1, D3 e2, double &w, double &u, double &v) {
    const D3 tmp1 = dcross(d, e2), t2r = dsub(o, p0);
    const double f = 1.0 / ddot(e1, tmp1);
    u = ddot(t2r, tmp1) * f;
    if (u < 0.0 || u > 1.0) return 1e20;
    const D3 tmp2 = dcross(t2r, e1);
    v = ddot(d, tmp2) * f;
    w = 1.0 - u - v;
    if (v < 0.0 || w < 0.0) return 1e20;
    const double t = ddot(e2, tmp2) * f;
    return t > 1e-5 ? t : 1e20;
}

// sphere_enter_t / sphere_exit_t in double.
RTG_HD double sphere_t64(D3 o, D3 d, D3 c, double rsq, bool exit) {
    const D3 s2r = dsub(o, c);
    const double b = ddot(d, s2r);
    if (!exit && b > 0) return 1e20;
    const double dis = b * b - ddot(s2r, s2r) + rsq;
    if (!exit && dis < 0) return 1e20;
    return exit ? -b + sqrt(dis) : -(b + sqrt(dis));
}

struct NearBand {
    static constexpr int kMax = 6;
    int n = 0;
    int pi[kMax];
    float t[kMax];
    RTG_HD void add(int p, float tt) {
        for (int k = 0; k < n; ++k)
            if (pi[k] == p) return;
        if (n < kMax) pi[n] = p, t[n] = tt, ++n;
    }
};

// FP64 re-test of primitive slot pi for query q; fills `h` when it hits.
RTG_HD bool retest64(const SceneView &S, const TraceRay &q, int pi, double &t64, Hit &h) {
    const D3 o = d3(q.r.o), d = d3(q.r.d);
    const float4 A = S.prims[3 * pi];
    const int tag = as_int(A.w);
    if (tag >= 0) {
        const float4 B = S.prims[3 * pi + 1];
        const int obj = as_int(B.w);
        if (tag == q.src.prim && obj == q.src.obj && !q.self_test) return false;
        const float4 *P = S.tri_pos + 3 * (size_t)tag;
        const D3 v0 = D3{P[0].x, P[0].y, P[0].z}, v1 = D3{P[1].x, P[1].y, P[1].z}, v2 = D3{P[2].x, P[2].y, P[2].z};
        const D3 p0 = dadd(v0, d3(S.objs[obj].pos));
        double w, u, v;
        t64 = tri_test64(o, d, p0, dsub(v1, v0), dsub(v2, v0), w, u, v);
        if (!(t64 < 1e19)) return false;
        h = Hit{(float)t64, HIT_TRI, obj, tag, (float)w, (float)u, (float)v};
        return true;
    }
    const int obj = ~tag;
    const bool exit = q.src.obj == obj;
    if (exit && !q.inside) return false;
    const DevObject &ob = S.objs[obj];
    t64 = sphere_t64(o, d, d3(ob.pos), (double)ob.radius * (double)ob.radius, exit);
    if (!(t64 > 1e-6 && t64 < 1e19)) return false;
    h = Hit{(float)t64, HIT_SPHERE, obj, -1, 0, 0, 0};
    return true;
}

// The three 16-B parts of BVH primitive slot pi, read together: one memory
// round trip per primitive test instead of a read of the tag followed by the
// reads its branch needs (a sphere leaves the third part unused).
struct PrimRec {
    float4 a, b, c;
};
RTG_HD PrimRec load_prim(const SceneView &S, int pi) {
    const float4 *p = S.prims + 3 * (size_t)pi;
    PrimRec r{p[0], p[1], p[2]};
#if defined(__HIP_DEVICE_COMPILE__)
    // uses every component here (no instruction), so the three reads are
    // issued as whole 16-B loads before the tag decides the branch
    __asm__ volatile("" ::"v"(r.a.x), "v"(r.a.y), "v"(r.a.z), "v"(r.a.w), "v"(r.b.x), "v"(r.b.y), "v"(r.b.z),
                     "v"(r.b.w), "v"(r.c.x), "v"(r.c.y), "v"(r.c.z), "v"(r.c.w));
#endif
    return r;
}

// Closest hit (RayTracer.cpp:455-465).  kRetest: FP64 near-tie re-test (see
// above); *retested counts the queries that were re-tested.  kBefore: only
// objects with index < `before` take part (see child_src).
template <bool kRetest = false, bool kBefore = false, class Stack>
RTG_HD Hit trace_closest(const SceneView &S, const TraceRay &q, Stack &st, int *retested = nullptr,
                         int before = 0x7fffffff) {
    Hit best{kMiss, HIT_NONE, -1, -1, 0, 0, 0};
    brute_objects(S, q, false, 0.0f, best, kBefore ? before : 0x7fffffff);
    uint64_t bkey = best.kind != HIT_NONE ? hit_key(best.obj, best.prim) : ~0ull;
    float tmax = best.t;
    int bpi = -1, pi2 = -1;  // primitive slots of the FP32 winner and runner-up (BVH primitives)
    float t2 = kMiss;
    NearBand band;
    auto take = [&](float t, uint64_t k, int pi, const Hit &h, float &tm) {
        if (t < best.t || (t == best.t && k < bkey)) {
            if (kRetest && bpi >= 0) t2 = best.t, pi2 = bpi;
            best = h;
            bkey = k;
            bpi = pi;
            tm = t;
        } else if (kRetest && t < t2) {
            t2 = t, pi2 = pi;
        }
    };
    bvh_walk<!kRetest>(S, q.r, tmax, st, [&](int pi, float &tm) -> bool {
        const PrimRec R = load_prim(S, pi);
        const float4 A = R.a;
        const int tag = as_int(A.w);
        const V3 p = V3{A.x, A.y, A.z};
        if (tag >= 0) {  // triangle, tag = global index
            const float4 B = R.b, C = R.c;
            const int obj = as_int(B.w);
            if (kBefore && obj >= before) return false;
            if (tag == q.src.prim && obj == q.src.obj && !q.self_test) return false;
            const V3 e1 = V3{B.x, B.y, B.z}, e2 = V3{C.x, C.y, C.z};
            float w, u, v, t;
            if constexpr (kRetest) {
                // tri_test's values without its early exits (same operations)
                const V3 tmp1 = cross(q.r.d, e2), t2r = q.r.o - p;
                const float f = 1.0f / dot(e1, tmp1);
                u = dot(t2r, tmp1) * f;
                const V3 tmp2 = cross(t2r, e1);
                v = dot(q.r.d, tmp2) * f;
                w = 1.0f - u - v;
                const float tt = dot(e2, tmp2) * f;
                if (tt == tt && tt > 0.0f) {
                    const float eps = 1e-6f;
                    const bool marginal = fabsf(u) <= eps || fabsf(u - 1.0f) <= eps || fabsf(v) <= eps ||
                                          fabsf(w) <= eps || fabsf(tt - 1e-5f) <= eps;
                    const bool inside = u >= -eps && u <= 1.0f + eps && v >= -eps && w >= -eps;
                    if (marginal && inside) band.add(pi, tt);
                }
                const bool ok = !(u < 0.0f || u > 1.0f) && !(v < 0.0f || w < 0.0f) && tt > 1e-5f;
                t = ok ? tt : kMiss;
            } else {
                t = tri_test(q.r, e1, e2, p, w, u, v);
            }
            if (t < kMiss) take(t, hit_key(obj, tag), pi, Hit{t, HIT_TRI, obj, tag, w, u, v}, tm);
        } else {  // sphere, tag = ~object
            const int obj = ~tag;
            if (kBefore && obj >= before) return false;
            const float4 B = R.b;
            float t;
            if (q.src.obj == obj) {
                if (!q.inside) return false;
                t = sphere_exit_t(q.r, p, B.x);
            } else {
                t = sphere_enter_t(q.r, p, B.x);
            }
            if constexpr (kRetest) {
                const V3 s2r = q.r.o - p;
                const float b = dot(q.r.d, s2r);
                const float dis = b * b - dot(s2r, s2r) + B.x;
                if (fabsf(dis) <= 1e-5f * (B.x + b * b)) band.add(pi, -b);
            }
            if (t > 1e-6f && t < kMiss) take(t, hit_key(obj, -1), pi, Hit{t, HIT_SPHERE, obj, -1, 0, 0, 0}, tm);
        }
        return false;
    });
    if constexpr (kRetest) {
        if (bpi >= 0 || best.kind == HIT_NONE) {
            const bool tie = bpi >= 0 && pi2 >= 0 && t2 < kMiss && t2 - best.t <= 4.0f * (nextafterf(best.t, kMiss) - best.t);
            const float lim = best.kind == HIT_NONE ? 1e30f : best.t * (1.0f + 1e-5f);
            int cand[NearBand::kMax + 2], nc = 0;
            for (int k = 0; k < band.n; ++k)
                if (band.t[k] <= lim) cand[nc++] = band.pi[k];
            if (tie || nc > 0) {
                if (bpi >= 0) cand[nc++] = bpi;
                if (tie) cand[nc++] = pi2;
                double bt = 1e300;
                uint64_t bk = ~0ull;
                Hit nb = best;
                for (int k = 0; k < nc; ++k) {
                    double t64;
                    Hit h;
                    if (!retest64(S, q, cand[k], t64, h)) continue;
                    const uint64_t key = hit_key(h.obj, h.prim);
                    if (t64 < bt || (t64 == bt && key < bk)) bt = t64, bk = key, nb = h;
                }
                if (bt < 1e19 && (nb.obj != best.obj || nb.prim != best.prim || nb.kind != best.kind)) best = nb;
                if (retested) ++*retested;
            }
        }
    }
    return best;
}

// Identity RTfrac hands to the children and shadow rays of a hit: its
// `newobj` (RayTracer.cpp:455-465), the last object in scan order whose hit
// changed HitRes::obj away from the ray's source identity.  That is the
// winner, except when the winner is the source object itself hit again from
// inside (a sphere or BallPlane-sphere exit): then it is the object that was
// the running closest hit before the source object's turn in the scan, i.e.
// the closest hit among the objects scanned before it, or the source identity
// when none of them was hit.  Returns obj == -2 when it is the winner.
RTG_HD bool is_self_rehit(const TraceRay &q, const Hit &h) {
    return q.src.obj >= 0 && h.obj == q.src.obj &&
           (h.kind == HIT_SPHERE || (h.kind == HIT_BALL && h.prim == q.src.prim));
}
// The identity from `b`, the closest hit among the objects before the source.
RTG_HD SrcId child_src_of(const SceneView &S, const TraceRay &q, const Hit &b) {
    if (b.kind == HIT_NONE) return q.src;
    SrcId id{b.obj, b.kind == HIT_TRI || b.kind == HIT_BALL ? b.prim : -1, -1};
    if (b.kind == HIT_TRI) id.bin = record_bin(S, q.r, q.src, b.prim);
    return id;
}
template <class Stack>
RTG_HD SrcId child_src(const SceneView &S, const TraceRay &q, const Hit &h, Stack &st) {
    if (!is_self_rehit(q, h)) return SrcId{-2, -1, -1};
    return child_src_of(S, q, trace_closest<false, true>(S, q, st, nullptr, h.obj));
}

// Any hit closer than `dis` (shadow test, RayTracer.cpp:510-520).
template <class Stack>
RTG_HD bool trace_any(const SceneView &S, const TraceRay &q, float dis, Stack &st) {
    Hit dummy{kMiss, HIT_NONE, -1, -1, 0, 0, 0};
    if (brute_objects(S, q, true, dis, dummy)) return true;
    bool blocked = false;
    float tmax = dis;
    bvh_walk<true, false>(S, q.r, tmax, st, [&](int pi, float &) -> bool {
        const PrimRec R = load_prim(S, pi);
        const float4 A = R.a;
        const int tag = as_int(A.w);
        const V3 p = V3{A.x, A.y, A.z};
        float t;
        if (tag >= 0) {
            const float4 B = R.b, C = R.c;
            const int obj = as_int(B.w);
            if (tag == q.src.prim && obj == q.src.obj && !q.self_test) return false;
            float w, u, v;
            t = tri_test(q.r, V3{B.x, B.y, B.z}, V3{C.x, C.y, C.z}, p, w, u, v);
        } else {
            const int obj = ~tag;
            const float4 B = R.b;
            if (q.src.obj == obj) {
                if (!q.inside) return false;
                blocked = B.y < dis;  // HitRes(radius) (Basic3DObject.cpp:143-144)
                return blocked;
            }
            t = sphere_enter_t(q.r, p, B.x);
            if (!(t > 1e-6f)) return false;
        }
        blocked = t < dis;
        return blocked;
    });
    return blocked;
}

}  // namespace rtg
