// rtg_wave.hip — wavefront engine: per-depth kernels over compacted queues.
//
// Level k of the ray tree (k = 0: primary rays, generated on the fly):
//   extend  closest hit of every queued ray                      -> hits[k]
//   shade   hit record + texel into a 64-byte node; child rays (reflection,
//           refraction) and shadow-ray entries appended to the next queues
//           with one atomic per wave (ballot/prefix-sum compaction)
//   occlude any-hit of every shadow entry; blocked lights set bits in the
//           node's light mask
// then, bottom-up from the deepest level:
//   combine local Blinn-Phong from the light mask, composed with the node's
//           children in the reference's order; written into the parent's
//           child slot (level 0: the sample colour)
//   resolve average of the samples of a pixel, quantised to RGB8 (with one
//           sample per pixel, level 0's combine writes the pixel itself)
// The per-node arithmetic is the megakernel's (shade_local, compose), so the
// two engines give bit-identical frames.
//
// Batches.  The kernels are templated on B.  B = false renders one request
// whose parameters and scene are kernel arguments.  B = true renders many
// requests, over many scenes, as one wavefront: level 0 concatenates every
// request's samples (each request's range is a whole number of 64-sample
// chunks, so a wave of primary rays belongs to one request); deeper queues
// mix requests, every ray and node carrying its request index.  A request's
// rays see only its own scene, so its frame is identical to rendering it
// alone.
#include <algorithm>
#include <cstring>

#include "rt_devstack.hpp"
#include "rt_prep.hpp"
#include "rtg_internal.hpp"

namespace rtg {

struct WRay {  // 48 B
    float ox, oy, oz, dx, dy, dz, bwc, mtlrfr;
    int link;  // 2 * parent node + slot (0 reflection, 1 refraction)
    int src_obj, src_prim;
    int flags;  // bits 0-2 type, bit 3 inside, bits 4-6 source copy bin, bits 8-31 request
};
struct WHit {  // 32 B
    float t, w, u, v;
    int kind, obj, prim;
    int self;  // 1: a self re-hit whose children take the identity stored in cs[i] (child_src)
};
struct WNode {  // 64 B
    float px, py, pz, nx, ny, nz, dx, dy, dz, vr, vg, vb, t;
    int mtl_req;  // bits 0-19 material, 20-31 request
    int id_prim;
    int bits;  // bits 0-23 object, 24-26 copy bin, 27 valid, 28 inside, 29 total internal reflection,
               // 30 a reflection child was traced, 31 a refraction child was traced
};
constexpr int kMtlBits = 20;
constexpr int kMaxBatch = 1 << (32 - kMtlBits);

// Device counters of one level, zeroed at the start of a frame; each in its
// own 256-byte line.
struct LevelCtr {
    LineCtr n_next;       // child rays appended for level k + 1
    LineCtr n_shadow;     // shadow-ray entries appended
    LineCtr sched[3][8];  // WaveQueue counters of extend, shade, occlude
};
constexpr int kCtrWords = (int)(sizeof(LevelCtr) / sizeof(unsigned));
// The counters of a pass: one LevelCtr per level, then the pass's abort word
// (non-zero once a kernel of the pass stopped for a cancel).
constexpr int kPassCtrWords = Workspace::kLevels * kCtrWords + (int)(sizeof(LineCtr) / sizeof(unsigned));

// One request's render parameters: a kernel argument (B = false) or an entry
// of the batch's request table (B = true).
struct WaveReq {
    int W, H, spp, max_level, tile_compat;
    int retest;       // RTG_FLAG_FP64_RETEST: FP64 near-tie re-test of closest hits
    int no_refract;   // the reflect test (RTflec, RayTracer.cpp:332-448): no refraction branch
    int tile_lg;      // pixel blocks of (1 << tile_lg) x (64 >> tile_lg): one wave's primary rays
    int bw;           // pixel blocks per row
    int row0;         // first row of pixel blocks this pass covers (a frame may take several passes)
    int n_samples;    // pixel blocks of the pass * 64 * spp
    int sample_base;  // first level-0 queue slot of the request
    int scene;        // index into the batch's scene table
    float dpf;
    uint64_t seed;
    const float *ax, *by;  // ray-generation tables (RayTracer.cpp:13-27)
    uint8_t *rgb;
    int32_t *ids;
    float *tprim;
};

struct WaveArgs {
    WaveReq r0;               // B = false
    const WaveReq *reqs;      // B = true
    const SceneView *scenes;  // B = true
    const int *chunk_req;     // B = true: request of each 64-sample chunk of level 0 (-1: padding)
    int n_samples;            // level-0 queue length
    int grab_lg;              // WaveQueue grabs of 64 << grab_lg items
    unsigned long long *counts;
    unsigned long long *retests;  // closest-hit queries re-tested in FP64
    int *spill;
    const int *cancel;  // mapped cancel flag of an asynchronous job, or null
    unsigned *abort;    // the pass's abort word
};

// Cooperative cancel: a wave that finds the job's cancel flag set marks the
// pass aborted and takes no work.  Called by every lane of a full wave before
// its work loop (wave-uniform result; no barrier follows it).
__device__ inline bool wave_cancelled(const WaveArgs &A) {
    if (A.cancel == nullptr) return false;
    int c = 0;
    if (__lane_id() == 0) c = __builtin_nontemporal_load(A.cancel);
    c = __builtin_amdgcn_readfirstlane(c);
    if (c != 0 && __lane_id() == 0) atomicOr(A.abort, 1u);
    return c != 0;
}

template <bool B>
__device__ inline const WaveReq &req_at(const WaveArgs &A, int r) {
    if constexpr (B)
        return A.reqs[r];
    else
        return A.r0;
}
template <bool B>
__device__ inline const SceneView &scene_at(const SceneView &S0, const WaveArgs &A, int sc) {
    if constexpr (B)
        return A.scenes[sc];
    else
        return S0;
}

__device__ inline bool sample_pixel(const WaveReq &R, int s, int &px, int &py, int &si) {
    const int per_block = 64 * R.spp;
    const int pb = s / per_block, r = s - pb * per_block;
    const int lane = r / R.spp;
    si = r - lane * R.spp;
    const int lg = R.tile_lg;
    px = (pb % R.bw) * (1 << lg) + (lane & ((1 << lg) - 1));
    py = (R.row0 + pb / R.bw) * (64 >> lg) + (lane >> lg);
    if (px >= R.W || py >= R.H) return false;
    if (R.tile_compat && (px >= ((R.W >> 6) << 6) || py >= ((R.H >> 6) << 6))) return false;
    return true;
}

// Primary ray of sample s of pixel (px, py): the reference's pixel-corner ray
// for spp 1 (RayTracer.cpp:26-27), 2x2 jittered strata otherwise.
__device__ inline TraceRay primary_ray_w(const SceneView &S, const WaveReq &R, int px, int py, int s) {
    float a = R.ax[px], b = R.by[py];
    if (R.spp > 1) {
        const uint64_t h = mix64(R.seed ^ ((uint64_t)(py * R.W + px) * 0x9E3779B97F4A7C15ull) ^
                                 ((uint64_t)s * 0xC2B2AE3D27D4EB4Full));
        const float jx = (float)(h >> 40) * (1.0f / 16777216.0f);
        const float jy = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        a = a + (((float)(s & 1) + jx) * 0.5f) * R.dpf;
        b = b + (((float)((s >> 1) & 1) + jy) * 0.5f) * R.dpf;
    }
    TraceRay q;
    q.r.o = S.cam.pos;
    q.r.d = normalize((S.cam.n + S.cam.u * a) + S.cam.v * b);
    q.type = RAY_BASE;
    q.inside = 0;
    q.src = no_src();
    q.self_test = 0;
    return q;
}

__device__ inline TraceRay load_ray(const WRay &w, float &bwc, float &mtlrfr) {
    TraceRay q;
    q.r.o = V3{w.ox, w.oy, w.oz};
    q.r.d = V3{w.dx, w.dy, w.dz};
    q.type = w.flags & 7;
    q.inside = (w.flags >> 3) & 1;
    q.src = SrcId{w.src_obj, w.src_prim, (w.flags >> 4) & 7};
    q.self_test = 0;
    bwc = w.bwc;
    mtlrfr = w.mtlrfr;
    return q;
}

__device__ inline WRay make_wray(const TraceRay &q, float bwc, float mtlrfr, int link, int req) {
    WRay w;
    w.ox = q.r.o.x, w.oy = q.r.o.y, w.oz = q.r.o.z;
    w.dx = q.r.d.x, w.dy = q.r.d.y, w.dz = q.r.d.z;
    w.bwc = bwc, w.mtlrfr = mtlrfr;
    w.link = link;
    w.src_obj = q.src.obj, w.src_prim = q.src.prim;
    w.flags = (q.type & 7) | ((q.inside & 1) << 3) | ((q.src.bin & 7) << 4) | (req << 8);
    return w;
}

__device__ inline int node_req(const WNode &nd) { return (int)((unsigned)nd.mtl_req >> kMtlBits); }
__device__ inline int node_mtl(const WNode &nd) { return nd.mtl_req & ((1 << kMtlBits) - 1); }
__device__ inline SrcId node_id(const WNode &nd) { return SrcId{nd.bits & 0xFFFFFF, nd.id_prim, (nd.bits >> 24) & 7}; }

__device__ inline void add_counts(unsigned long long *counts, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    a = wave_sum_u32(a), b = wave_sum_u32(b), c = wave_sum_u32(c), d = wave_sum_u32(d);
    if (__lane_id() == 0) {
        if (a) atomicAdd(&counts[0], (unsigned long long)a);
        if (b) atomicAdd(&counts[1], (unsigned long long)b);
        if (c) atomicAdd(&counts[2], (unsigned long long)c);
        if (d) atomicAdd(&counts[3], (unsigned long long)d);
    }
}

__device__ inline uint32_t wave_or_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    return v;
}

// Request of level-0 sample i (-1: padding).
template <bool B>
__device__ inline int sample_req(const WaveArgs &A, int i) {
    if constexpr (B)
        return A.chunk_req[i >> 6];
    else
        return 0;
}

// A persistent traversal grid is sized for the level's bound, and the deep
// levels hold far fewer items: a block whose waves would each start past the
// last 64-item chunk finds no work, and leaves before copying the top nodes
// into LDS (block-uniform, so ahead of the barrier there).  The blocks that
// stay take every chunk: waves keep grabbing until the queue is empty.
__device__ inline bool surplus_block(int n) {
    return (int)blockIdx.x * (kBlockThreads / 64) >= ((n + 63) >> 6);
}

// Grab size (64 << lg items) for n items: the tuning's, doubled up to eight
// chunks while every wave of the grid still gets four grabs or more.  Each
// grab is a returning atomic on its XCD's counter, a round trip of about 2 us
// when every wave contends; small levels keep single chunks, so that their
// items still spread over the grid.
__device__ inline int grab_lg_for(int n, int lg) {
    const int waves = (int)gridDim.x * (kBlockThreads / 64), chunks = (n + 63) >> 6;
    while (lg < 3 && (chunks >> (lg + 1)) >= 4 * waves) ++lg;
    return lg;
}

// R: some request asks for the FP64 near-tie re-test (a separate kernel, so
// the plain closest-hit walk keeps its registers).
template <bool B, bool R>
__global__ __launch_bounds__(kBlockThreads, 6) void k_extend(SceneView S0, WaveArgs A, int k, const WRay *rays,
                                                             const unsigned *n_ptr, WHit *hits, int4 *cs,
                                                             LineCtr *sched) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    if (surplus_block(n)) return;
    __shared__ int lds[kLdsStack * kBlockThreads];
    DevStack st = make_stack(lds, A.spill);
    if constexpr (!B) {  // one scene: its top nodes in LDS
        __shared__ Bvh4Node top_nodes[kTopNodes];
        stack_top_nodes(st, top_nodes, S0);
    }
    if (wave_cancelled(A)) return;
    WaveQueue wq(sched, n, A.grab_lg);
    uint32_t c_prim = 0, c_refl = 0, c_refr = 0;
    int n_retest = 0;
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int i = base + __lane_id();
        if (i >= n) continue;
        int r = 0;
        WRay w{};
        if (k == 0) {
            r = sample_req<B>(A, i);
        } else {
            w = rays[i];
            r = w.flags >> 8;
        }
        const WaveReq &Rq = req_at<B>(A, r < 0 ? 0 : r);
        const SceneView &S = scene_at<B>(S0, A, Rq.scene);
        TraceRay q;
        int px = 0, py = 0, si = 0;
        bool valid = r >= 0;
        if (k == 0) {
            if (valid) valid = sample_pixel(Rq, i - Rq.sample_base, px, py, si);
            if (valid) q = primary_ray_w(S, Rq, px, py, si);
        } else {
            float bwc, mtlrfr;
            q = load_ray(w, bwc, mtlrfr);
            if (q.src.obj >= 0) q.self_test = self_testable(S, q.r, q.src, kMiss);
        }
        // Pass 0 is the closest hit.  A self re-hit (child_src) takes a second
        // pass over the objects before the source object; the plain kernel
        // runs it through the same walk (one inlined copy of the traversal),
        // after the hit record is stored so that little state stays live.
        int before = 0x7fffffff;
        for (;;) {
            Hit h{kMiss, HIT_NONE, -1, -1, 0, 0, 0};
            if (valid) {
                if constexpr (R)
                    h = Rq.retest ? trace_closest<true>(S, q, st, &n_retest) : trace_closest<false>(S, q, st);
                else
                    h = trace_closest<false, true>(S, q, st, nullptr, before);
            }
            if (before != 0x7fffffff) {
                const SrcId c = child_src_of(S, q, h);
                cs[i] = int4{c.obj, c.prim, c.bin, 0};
                break;
            }
            if (valid) {
                if (k == 0)
                    ++c_prim;
                else if (q.type == RAY_REFLECT)
                    ++c_refl;
                else
                    ++c_refr;
            }
            const bool self = valid && h.kind != HIT_NONE && is_self_rehit(q, h);
            hits[i] = WHit{h.t, h.w, h.u, h.v, h.kind, h.obj, h.prim, self ? 1 : 0};
            if (!self) break;
            if constexpr (R) {
                const SrcId c = child_src(S, q, h, st);
                cs[i] = int4{c.obj, c.prim, c.bin, 0};
                break;
            }
            before = h.obj;
        }
    }
    add_counts(A.counts, c_prim, 0, c_refl, c_refr);
    if constexpr (R) {
        const uint32_t nr = wave_sum_u32((uint32_t)n_retest);
        if (__lane_id() == 0 && nr) atomicAdd(A.retests, (unsigned long long)nr);
    }
}

// k_shade's appends are staged per wave in LDS and reach their queue a buffer
// at a time, one atomic each: returning atomics on one queue counter
// serialise across the chip (about 11 ns apart on MI355X), and two per 64
// items bounded the kernel.
constexpr int kStageRays = 160;    // child rays (48 B each)
constexpr int kStageShadow = 512;  // shadow-ray entries

// Orders a wave's LDS writes before its later LDS reads by other lanes (the
// LDS executes one wave's accesses in order; this stops the compiler moving
// them).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix of `count` over the wave; `total` (wave-uniform) gets the sum.
__device__ inline int wave_prefix(int count, int &total) {
    const int lane = __lane_id();
    int x = count;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    total = __builtin_amdgcn_readfirstlane(__shfl(x, 63, 64));
    return x - count;
}

// Move the wave's n staged items (n wave-uniform) to the end of a queue.
template <class T>
__device__ inline void flush_stage(const T *stage, int n, unsigned *counter, T *queue) {
    if (n == 0) return;
    int b = 0;
    if (__lane_id() == 0) b = (int)atomicAdd(counter, (unsigned)n);
    b = __builtin_amdgcn_readfirstlane(b);
    wave_lds_sync();
    for (int j = __lane_id(); j < n; j += 64) queue[b + j] = stage[j];
    wave_lds_sync();  // before the stage is written again
}

template <bool B>
__global__ __launch_bounds__(kBlockThreads, 4) void k_shade(SceneView S0, WaveArgs A, int k, const WRay *rays,
                                                         const unsigned *n_ptr, const WHit *hits, const int4 *cs,
                                                         WNode *nodes,
                                                         unsigned *masks, WRay *next, LevelCtr *ctr,
                                                         unsigned *shadow, float4 *slots) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    if (wave_cancelled(A)) return;
    WaveQueue wq(ctr->sched[1], n, grab_lg_for(n, A.grab_lg));
    uint32_t c_shadow = 0;
    __shared__ __align__(16) WRay stage_rays[kBlockThreads / 64][kStageRays];
    __shared__ unsigned stage_shadow[kBlockThreads / 64][kStageShadow];
    WRay *const sr = stage_rays[threadIdx.x >> 6];
    unsigned *const ss = stage_shadow[threadIdx.x >> 6];
    int n_sr = 0, n_ss = 0;  // staged items (wave-uniform)
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int i = base + __lane_id();
        int n_child = 0;
        bool lit = false;
        WRay c0, c1;  // children in append order (registers, no indexed array)
        V3 P = V3{0, 0, 0}, N = V3{0, 0, 0}, D = V3{0, 0, 0};
        uint32_t on = 0;  // switched-on lights of the ray's scene
        const SceneView *Sp = &S0;
        if (i < n) {
            int r = 0;
            WRay w{};
            if (k == 0) {
                r = sample_req<B>(A, i);
            } else {
                w = rays[i];
                r = w.flags >> 8;
            }
            const WaveReq &R = req_at<B>(A, r < 0 ? 0 : r);
            const SceneView &S = scene_at<B>(S0, A, R.scene);
            Sp = &S;
            TraceRay q;
            float bwc = 1.0f, mtlrfr = 1.0f;
            bool valid = r >= 0;
            if (k == 0) {
                int px, py, si;
                if (valid) valid = sample_pixel(R, i - R.sample_base, px, py, si);
                if (valid) q = primary_ray_w(S, R, px, py, si);
                if (valid && si == 0 && (R.ids || R.tprim)) {  // primary hit of the pixel's first sample
                    const WHit wp = hits[i];
                    const size_t pix = (size_t)py * R.W + px;
                    if (R.ids) {
                        R.ids[2 * pix] = wp.obj;
                        R.ids[2 * pix + 1] = id_prim(S, Hit{wp.t, wp.kind, wp.obj, wp.prim, wp.w, wp.u, wp.v});
                    }
                    if (R.tprim) R.tprim[pix] = wp.t;
                }
            } else {
                q = load_ray(w, bwc, mtlrfr);
            }
            const WHit wh = hits[i];
            const float zNear = k == 0 ? S.cam.zNear : 0.0f;
            WNode nd;
            nd.bits = 0;
            if (valid && wh.kind != HIT_NONE && !(wh.t > S.cam.zFar || wh.t < zNear)) {
                const Hit h{wh.t, wh.kind, wh.obj, wh.prim, wh.w, wh.u, wh.v};
                HitRec hr = make_hitrec(S, q, h);
                if (wh.self) {
                    const int4 c = cs[i];
                    hr.id = SrcId{c.x, c.y, c.z};
                }
                const V3 vc = texel(S, hr.tex, hr.tu, hr.tv);
                const DevMaterial &m = S.mtls[hr.mtl];
                const float m_refract = R.no_refract ? 0.0f : m.refract;
                int tir = 0, kids = 0;  // bit 0: reflection child, bit 1: refraction child
                if (m.reflect > 0.01f && k + 1 <= R.max_level && bwc * m.reflect >= 1e-5f) {
                    c0 = make_wray(reflect_ray(q, hr), bwc * m.reflect, 1.0f, 2 * i, r);
                    n_child = 1;
                    kids = 1;
                }
                if (m_refract > 0.01f) {
                    TraceRay cq;
                    if (!refract_ray(q, mtlrfr, hr, cq)) {
                        tir = 1;
                    } else if (k + 1 <= R.max_level && bwc * m_refract >= 1e-5f) {
                        const WRay cw = make_wray(cq, bwc * m_refract, hr.rfr, 2 * i + 1, r);
                        if (n_child == 0)
                            c0 = cw;
                        else
                            c1 = cw;
                        ++n_child;
                        kids |= 2;
                    }
                }
                nd.px = hr.P.x, nd.py = hr.P.y, nd.pz = hr.P.z;
                nd.nx = hr.N.x, nd.ny = hr.N.y, nd.nz = hr.N.z;
                nd.dx = q.r.d.x, nd.dy = q.r.d.y, nd.dz = q.r.d.z;
                nd.vr = vc.x, nd.vg = vc.y, nd.vb = vc.z;
                nd.t = hr.t;
                nd.mtl_req = hr.mtl | (r << kMtlBits);
                nd.id_prim = hr.id.prim;
                nd.bits = (hr.id.obj & 0xFFFFFF) | ((hr.id.bin & 7) << 24) | (1 << 27) | ((hr.inside & 1) << 28) |
                          (tir << 29) | (int)((unsigned)kids << 30);
                lit = true;
                P = hr.P, N = hr.N, D = q.r.d;
                on = S.lights_on;
                c_shadow += __builtin_popcount(on);  // the reference traces one shadow ray per switched-on light
            }
            // a ray without a hit leaves only its node's bits (nothing reads more:
            // its light mask and child slots neither)
            if ((nd.bits >> 27) & 1) {
                nodes[i] = nd;
                masks[i] = 0;
            } else {
                nodes[i].bits = 0;
            }
            // k_combine of the next level writes the child slot of every
            // child traced (bits 30, 31); k_combine of this level reads only
            // those, so the slots are not cleared here
        }
        int t_child;
        const int pc = wave_prefix(n_child, t_child);
        if (n_sr + t_child > kStageRays) {
            flush_stage(sr, n_sr, &ctr->n_next.v, next);
            n_sr = 0;
        }
        if (n_child > 0) sr[n_sr + pc] = c0;
        if (n_child > 1) sr[n_sr + pc + 1] = c1;
        n_sr += t_child;
        // Shadow entries (node << 5 | light), light-major within the wave's
        // block, so that a wave of k_occlude traces neighbouring points toward
        // one light.  Lights that cannot change the colour are not traced
        // (shade_local ignores their mask bit).
        uint32_t emit = 0;
        if (lit)
            for (uint32_t lm = on; lm; lm &= lm - 1) {
                const int li = __builtin_ctz(lm);
                if (light_dots(N, D, light_term(Sp->lights[li], P)).matters()) emit |= 1u << li;
            }
        const uint32_t any = B ? wave_or_u32(emit) : S0.lights_on;
        int total = 0;
        for (uint32_t lm = any; lm; lm &= lm - 1) total += __popcll(__ballot((emit >> __builtin_ctz(lm)) & 1u));
        total = __builtin_amdgcn_readfirstlane(total);
        const unsigned long long below = (1ull << __lane_id()) - 1ull;
        if (total > kStageShadow) {  // many lights: straight to the queue
            int sb = 0;
            if (__lane_id() == 0) sb = (int)atomicAdd(&ctr->n_shadow.v, (unsigned)total);
            sb = __builtin_amdgcn_readfirstlane(sb);
            for (uint32_t lm = any; lm; lm &= lm - 1) {
                const int li = __builtin_ctz(lm);
                const unsigned long long m = __ballot((emit >> li) & 1u);
                if ((emit >> li) & 1u) shadow[sb + __popcll(m & below)] = ((unsigned)i << 5) | (unsigned)li;
                sb += __popcll(m);
            }
        } else {
            if (n_ss + total > kStageShadow) {
                flush_stage(ss, n_ss, &ctr->n_shadow.v, shadow);
                n_ss = 0;
            }
            int sb = n_ss;
            for (uint32_t lm = any; lm; lm &= lm - 1) {
                const int li = __builtin_ctz(lm);
                const unsigned long long m = __ballot((emit >> li) & 1u);
                if ((emit >> li) & 1u) ss[sb + __popcll(m & below)] = ((unsigned)i << 5) | (unsigned)li;
                sb += __popcll(m);
            }
            n_ss += total;
        }
    }
    flush_stage(sr, n_sr, &ctr->n_next.v, next);
    flush_stage(ss, n_ss, &ctr->n_shadow.v, shadow);
    add_counts(A.counts, 0, c_shadow, 0, 0);
}

template <bool B>
__global__ __launch_bounds__(kBlockThreads, 8) void k_occlude(SceneView S0, WaveArgs A, const WNode *nodes,
                                                           const unsigned *shadow, LevelCtr *ctr, unsigned *masks) {
    const int n = (int)ctr->n_shadow.v;
    if (surplus_block(n)) return;
    __shared__ int lds[kLdsStack * kBlockThreads];
    DevStack st = make_stack(lds, A.spill);
    if constexpr (!B) {  // one scene: its top nodes in LDS
        __shared__ Bvh4Node top_nodes[kTopNodes];
        stack_top_nodes(st, top_nodes, S0);
    }
    if (wave_cancelled(A)) return;
    WaveQueue wq(ctr->sched[2], n, A.grab_lg);
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int j = base + __lane_id();
        if (j >= n) continue;
        const unsigned e = shadow[j];
        const int i = (int)(e >> 5), li = (int)(e & 31);
        const WNode nd = nodes[i];
        const SceneView &S = scene_at<B>(S0, A, req_at<B>(A, node_req(nd)).scene);
        const V3 P = V3{nd.px, nd.py, nd.pz};
        const LightTerm lt = light_term(S.lights[li], P);
        TraceRay sr;
        sr.r.o = P;
        sr.r.d = lt.p2l;
        sr.type = RAY_SHADOW;
        sr.inside = 0;
        sr.src = node_id(nd);
        sr.self_test = self_testable(S, sr.r, sr.src, lt.dis);
        if (trace_any(S, sr, lt.dis, st)) atomicOr(&masks[i], 1u << li);
    }
}

template <bool B>
__global__ __launch_bounds__(kBlockThreads, 8) void k_local(SceneView S0, WaveArgs A, int k, const unsigned *n_ptr,
                                                         WNode *nodes, const unsigned *masks) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    if (wave_cancelled(A)) return;
    const ChunkRange cr = xcd_chunks((n + kBlockThreads - 1) / kBlockThreads);
    for (int c = cr.first; c < cr.end; c += cr.step) {
        const int i = c * kBlockThreads + threadIdx.x;
        if (i >= n || !((nodes[i].bits >> 27) & 1)) continue;
        const WNode nd = nodes[i];
        const WaveReq &R = req_at<B>(A, node_req(nd));
        const SceneView &S = scene_at<B>(S0, A, R.scene);
        const unsigned mask = masks[i];
        const V3 L = shade_local(S, V3{nd.dx, nd.dy, nd.dz}, V3{nd.px, nd.py, nd.pz}, V3{nd.nx, nd.ny, nd.nz},
                                 node_mtl(nd), V3{nd.vr, nd.vg, nd.vb},
                                 [&](int li, const LightTerm &, bool) -> bool { return (mask >> li) & 1u; });
        nodes[i].vr = L.x, nodes[i].vg = L.y, nodes[i].vb = L.z;
    }
}

// local: the node's colour fields hold its local lighting already (k_local).
template <bool B>
__global__ __launch_bounds__(kBlockThreads, 8) void k_combine(SceneView S0, WaveArgs A, int k, const WRay *rays,
                                                           const unsigned *n_ptr, const WNode *nodes,
                                                           const unsigned *masks, const float4 *slots,
                                                           float4 *parent_slots, float4 *samples, int local) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    // Level 0 writes the pixels of requests with one sample per pixel (the
    // box filter of one sample is that sample, quantised), in place of
    // k_resolve.  Like k_resolve it reads the pass's abort word, which no
    // kernel sets any more once level 0 combines, so a pass a cancel stopped
    // keeps all its pixels untouched.
    if (k == 0) {
        if (A.cancel != nullptr && *(volatile unsigned *)A.abort != 0) return;
    } else if (wave_cancelled(A)) {
        return;
    }
    const ChunkRange cr = xcd_chunks((n + kBlockThreads - 1) / kBlockThreads);
    for (int c = cr.first; c < cr.end; c += cr.step) {
        const int i = c * kBlockThreads + threadIdx.x;
        if (i >= n) continue;
        float4 out = make_float4(0.0f, 0.0f, 0.0f, kMiss);
        if ((nodes[i].bits >> 27) & 1) {  // a node without a hit holds only its bits
            const WNode nd = nodes[i];
            const WaveReq &R = req_at<B>(A, node_req(nd));
            const SceneView &S = scene_at<B>(S0, A, R.scene);
            const int mtl = node_mtl(nd);
            V3 L = V3{nd.vr, nd.vg, nd.vb};
            if (!local) {
                const unsigned mask = masks[i];
                L = shade_local(S, V3{nd.dx, nd.dy, nd.dz}, V3{nd.px, nd.py, nd.pz}, V3{nd.nx, nd.ny, nd.nz}, mtl,
                                L, [&](int li, const LightTerm &, bool) -> bool { return (mask >> li) & 1u; });
            }
            const DevMaterial &m = S.mtls[mtl];
            const float4 none = make_float4(0.0f, 0.0f, 0.0f, kMiss);  // no child: black, no Beer factor
            const float4 a = (nd.bits >> 30) & 1 ? slots[2 * i] : none;
            const float4 b = (nd.bits >> 31) & 1 ? slots[2 * i + 1] : none;
            const V3 col = compose(L, m.reflect, R.no_refract ? 0.0f : m.refract, (nd.bits >> 29) & 1,
                                   (nd.bits >> 28) & 1, m.kd, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, b.w);
            out = make_float4(col.x, col.y, col.z, nd.t);
        }
        if (k != 0) {
            parent_slots[rays[i].link] = out;
            continue;
        }
        const int r = sample_req<B>(A, i);
        const WaveReq &Rq = req_at<B>(A, r < 0 ? 0 : r);
        if (Rq.spp != 1) {
            samples[i] = out;
            continue;
        }
        int px, py, si;
        if (r >= 0 && sample_pixel(Rq, i - Rq.sample_base, px, py, si)) {
            uint8_t *p = Rq.rgb + 3 * ((size_t)py * Rq.W + px);
            p[0] = quantize(out.x), p[1] = quantize(out.y), p[2] = quantize(out.z);
        }
    }
}

// Box filter of a request's samples, quantised (RayTracer.cpp:34-37,
// 3DElement.cpp:463-468).  Grid y = request.
template <bool B>
__global__ void k_resolve(WaveArgs A, const float4 *samples) {
    // a pass that a cancel stopped part-way keeps its pixels untouched (127)
    if (A.cancel != nullptr && *(volatile unsigned *)A.abort != 0) return;
    const WaveReq &R = req_at<B>(A, blockIdx.y);
    if (R.spp == 1) return;  // written by level 0's k_combine
    // pixels of the pass: block rows row0 .. row0 + n_samples / (bw 64 spp)
    const int lg = R.tile_lg, th_lg = 6 - lg;
    const int y0 = R.row0 << th_lg;
    const int y1 = R.bw > 0 ? min(R.H, y0 + ((R.n_samples / (R.bw * 64 * R.spp)) << th_lg)) : y0;
    const int npix = R.W * (y1 - y0);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npix; q += gridDim.x * blockDim.x) {
        const int p = y0 * R.W + q;
        const int px = p % R.W, py = p / R.W;
        if (R.tile_compat && (px >= ((R.W >> 6) << 6) || py >= ((R.H >> 6) << 6))) continue;
        const int pb = ((py >> th_lg) - R.row0) * R.bw + (px >> lg);
        const int lane = ((py & ((1 << th_lg) - 1)) << lg) | (px & ((1 << lg) - 1));
        const int s0 = R.sample_base + (pb * 64 + lane) * R.spp;
        V3 acc = V3{0, 0, 0};
        for (int s = 0; s < R.spp; ++s) {
            const float4 c = samples[s0 + s];
            acc = acc + V3{c.x, c.y, c.z};
        }
        if (R.spp > 1) acc = acc * (1.0f / (float)R.spp);
        R.rgb[3 * (size_t)p + 0] = quantize(acc.x);
        R.rgb[3 * (size_t)p + 1] = quantize(acc.y);
        R.rgb[3 * (size_t)p + 2] = quantize(acc.z);
    }
}

__global__ void k_zero_u32(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Batch: request of every 64-sample chunk of level 0 (binary search over the
// requests' sample ranges; -1 for padding chunks).
__global__ void k_chunk_req(const WaveReq *reqs, int n_reqs, int *chunk_req, int n_chunks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    const int s = c << 6;
    int lo = 0, hi = n_reqs - 1, r = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (reqs[mid].sample_base <= s) {
            r = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (r >= 0 && s >= reqs[r].sample_base + reqs[r].n_samples) r = -1;
    chunk_req[c] = r;
}

// Single frames with more samples run level by level (see launch_wavefront):
// 4M covers 1080p at 2 spp; the persistent kernels read each level's real
// count on the device, so a no-sync frame only pays for its queue memory.
constexpr int kNoSyncMaxSamples = 1 << 22;
// Level-0 samples per pass of a frame whose rays branch (see launch_wavefront).
constexpr int kPassSamples = 1 << 21;

// Resident blocks per CU of a kernel at kBlockThreads threads (registers and
// LDS permitting).
template <class K>
static int blocks_per_cu(K kernel) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlockThreads, 0) != hipSuccess || nb < 1) nb = 4;
    return nb;
}

template <bool B>
struct Grids {
    int extend, shade, occlude, occlude_alone;
    // pct: percent of the resident blocks (at least one block per CU).
    // k_shade takes half that share: its waves mostly wait on their appends
    // and stores, and fewer of them finish sooner (measured).  Shadow rays
    // that run beside no traversal of a next level (the last level's, or all
    // of them on one stream) take every resident block.
    explicit Grids(int cus, int pct = 100) {
        static const int oe = blocks_per_cu(k_extend<B, false>), os = blocks_per_cu(k_shade<B>),
                         oo = blocks_per_cu(k_occlude<B>);
        auto part = [&](int o, int p) { return std::max(cus, cus * o * p / 100); };
        extend = part(oe, pct), shade = part(os, pct / 2), occlude = part(oo, pct), occlude_alone = part(oo, 100);
    }
};

// Shape of a wavefront: what bounds its queues.
struct WavePlan {
    int n_samples;  // level-0 queue length
    int max_level;  // deepest level any request traces
    int fanout;     // most child rays per hit
    int n_on;       // most switched-on lights of a scene
    int n_reqs;     // grid y of k_resolve
    size_t bound[Workspace::kLevels];
    bool no_sync;
};

static size_t per_ray_bytes(int n_on) {
    return sizeof(WRay) + sizeof(WHit) + sizeof(int4) + sizeof(WNode) + sizeof(unsigned) + 2 * sizeof(float4) +
           (size_t)n_on * sizeof(unsigned);
}

size_t batch_queue_bytes(size_t samples, int max_level, int fanout, int n_on, size_t budget) {
    size_t need = 0, b = samples;
    const size_t cap = level_ray_cap(n_on);
    for (int k = 0; k <= max_level && b > 0; ++k) {
        if (b > cap) return 0;
        need += b * per_ray_bytes(n_on);
        if (need > budget) return 0;
        b *= (size_t)fanout;
    }
    return need;
}

static WavePlan make_plan(int n_samples, int max_level, int fanout, int n_on, int n_reqs, size_t budget) {
    WavePlan pl{};
    pl.n_samples = n_samples, pl.max_level = max_level, pl.fanout = fanout, pl.n_on = n_on, pl.n_reqs = n_reqs;
    const size_t per_ray = per_ray_bytes(n_on), cap = level_ray_cap(n_on);
    size_t need = 0;
    pl.no_sync = true;
    for (int k = 0; k <= max_level; ++k) {
        pl.bound[k] = k == 0 ? (size_t)n_samples : std::min<size_t>(pl.bound[k - 1] * (size_t)fanout, size_t(1) << 40);
        need += pl.bound[k] * per_ray;
        // a bound past the level cap: the levels are sized from their real
        // counts instead, each checked against the cap before it runs
        if (need > budget || pl.bound[k] > cap) pl.no_sync = false;
    }
    return pl;
}

// Grow the level-k buffers for n_k rays.
static int reserve_level(Workspace &ws, const WavePlan &pl, int k, size_t n_k) {
    const size_t cap_next = k < pl.max_level ? (size_t)std::max(1, pl.fanout) * n_k : 1;
    const int kn = k + 1 < Workspace::kLevels ? k + 1 : k;
    if (ws.grow(ws.w_hits[k], n_k * sizeof(WHit)) || ws.grow(ws.w_cs[k], n_k * sizeof(int4)) ||
        ws.grow(ws.w_nodes[k], n_k * sizeof(WNode)) ||
        ws.grow(ws.w_masks[k], n_k * sizeof(unsigned)) || ws.grow(ws.w_slots[k], 2 * n_k * sizeof(float4)) ||
        ws.grow(ws.w_shadow[k], std::max<size_t>(1, n_k * pl.n_on) * sizeof(unsigned)) ||
        ws.grow(ws.w_rays[kn], cap_next * sizeof(WRay)))
        return fail(RTG_E_NOMEM, "wavefront queues");
    return RTG_OK;
}

// Buffers of a no-sync plan, all sized up front (a captured graph must not
// see them move).
static int reserve_all(Workspace &ws, const WavePlan &pl, int cus) {
    const Grids<false> g0(cus);
    const Grids<true> g1(cus);
    const int gmax = std::max(std::max(g0.extend, g0.occlude_alone), std::max(g1.extend, g1.occlude_alone));
    if (ws.grow(ws.spill, spill_ints(gmax) * sizeof(int))) return fail(RTG_E_NOMEM, "spill");
    if (ws.grow(ws.spill2, spill_ints(std::max(g0.occlude, g1.occlude)) * sizeof(int)))
        return fail(RTG_E_NOMEM, "spill");
    if (ws.grow(ws.w_qcount, (size_t)kPassCtrWords * sizeof(unsigned))) return fail(RTG_E_NOMEM, "counters");
    if (ws.grow(ws.w_samples, (size_t)pl.n_samples * sizeof(float4))) return fail(RTG_E_NOMEM, "samples");
    for (int k = 0; k <= pl.max_level && pl.bound[k] > 0; ++k) {
        const int st = reserve_level(ws, pl, k, pl.bound[k]);
        if (st != RTG_OK) return st;
    }
    return RTG_OK;
}

// Enqueue the levels of a wavefront on `stream`.  With pl.no_sync every
// buffer must already hold its bound (reserve_all) and nothing here
// allocates or waits.
template <bool B>
static int enqueue_wave(const SceneView &S0, const WaveArgs &a, const WavePlan &pl, hipStream_t stream,
                        Workspace &ws, Arena &A, const Cancel *cancel, bool retest, bool overlap) {
    const int *cancel_host = cancel ? cancel->host : nullptr;
    const int cus = A.cus > 0 ? A.cus : 256;
    const Grids<B> g(cus, A.tune.grid_pct);
    if (!pl.no_sync) {
        if (ws.grow(ws.spill, spill_ints(std::max(g.extend, g.occlude_alone)) * sizeof(int)))
            return fail(RTG_E_NOMEM, "spill");
        if (ws.grow(ws.w_qcount, (size_t)kPassCtrWords * sizeof(unsigned)))
            return fail(RTG_E_NOMEM, "queue counters");
    }
    // The shadow rays of level k (k_occlude) depend only on
    // k_shade(k), and level k + 1 does not depend on them: with `overlap`
    // they run on the workspace's side stream, concurrently with the next
    // level, and join the render stream before k_combine (the last level's
    // stay on the render stream, see below).
    const bool ov = overlap && pl.n_on > 0;
    if (ov) {
        int st = ws.side_init();
        if (st != RTG_OK) return st;
        if (ws.grow(ws.spill2, spill_ints(g.occlude) * sizeof(int))) return fail(RTG_E_NOMEM, "spill");
    }
    hipStream_t so = ov ? ws.side : stream;
    // A large single frame on a caller's stream, all levels enqueued at once,
    // alternates the shadow rays of its levels k < max between the side
    // stream and the arena's own stream (idle meanwhile: the caller holds the
    // arena lock), so that level k's need not wait for level k - 1's; each
    // stream has its own spill area.  Smaller frames, and frames that read
    // level counts back, were slower with it (measured): their shadow rays
    // do not bound the frame.
    const bool ov2 = !B && ov && pl.no_sync && pl.n_samples >= (1 << 20) && A.stream != nullptr &&
                     A.stream != stream && A.stream != ws.side;
    if (ov2 && ws.grow(ws.spill3, spill_ints(g.occlude) * sizeof(int))) return fail(RTG_E_NOMEM, "spill");
    hipStream_t so2 = ov2 ? A.stream : so;
    bool forked = false, forked2 = false;  // a side stream took work (the join waits for it)
    auto join = [&]() -> int {
        if (forked) {
            RTG_HIP(hipEventRecord(ws.ev_join, so));
            RTG_HIP(hipStreamWaitEvent(stream, ws.ev_join, 0));
        }
        if (forked2) {
            RTG_HIP(hipEventRecord(ws.ev_join2, so2));
            RTG_HIP(hipStreamWaitEvent(stream, ws.ev_join2, 0));
        }
        return RTG_OK;
    };
    LevelCtr *ctr = ws.w_qcount.as<LevelCtr>();
    k_zero_u32<<<(kPassCtrWords + 255) / 256, 256, 0, stream>>>(ws.w_qcount.as<unsigned>(), kPassCtrWords);
    const int lblocks = std::max(8, cus * 16);  // k_local, k_combine
    const bool early_local = !B && pl.no_sync;
    auto grid = [](int gmax, size_t items) {
        return (int)std::max<size_t>(8, std::min<size_t>(gmax, (items + 127) / 128));
    };
    WaveArgs args = a;
    args.spill = ws.spill.as<int>();
    args.cancel = cancel ? cancel->dev : nullptr;
    args.abort = ws.w_qcount.as<unsigned>() + Workspace::kLevels * kCtrWords;
    WaveArgs oargs = args, oargs2 = args;
    if (ov) oargs.spill = ws.spill2.as<int>();
    if (ov2) oargs2.spill = ws.spill3.as<int>();
    size_t n_k = (size_t)pl.n_samples;
    int last = -1;
    const size_t cap = level_ray_cap(pl.n_on);
    for (int k = 0; k <= pl.max_level && n_k > 0; ++k) {
        if (cancel_host && *cancel_host) return join();
        if (n_k > cap) {
            (void)join();
            return fail(RTG_E_LIMIT, "wavefront level " + std::to_string(k) + " holds " + std::to_string(n_k) +
                                         " rays, more than the per-level limit of " + std::to_string(cap));
        }
        if (!pl.no_sync) {
            const int st = reserve_level(ws, pl, k, n_k);
            if (st != RTG_OK) {
                (void)join();
                return st;
            }
        }
        const WRay *rays = k == 0 ? nullptr : ws.w_rays[k].as<WRay>();
        const unsigned *n_ptr = k == 0 ? nullptr : &ctr[k - 1].n_next.v;
        WRay *next = ws.w_rays[k + 1 < Workspace::kLevels ? k + 1 : k].as<WRay>();
        int kt = ktime_begin(A, stream, KT_EXTEND, n_k);
        if (retest)
            k_extend<B, true><<<grid(g.extend, n_k), kBlockThreads, 0, stream>>>(
                S0, args, k, rays, n_ptr, ws.w_hits[k].as<WHit>(), ws.w_cs[k].as<int4>(), ctr[k].sched[0]);
        else
            k_extend<B, false><<<grid(g.extend, n_k), kBlockThreads, 0, stream>>>(
                S0, args, k, rays, n_ptr, ws.w_hits[k].as<WHit>(), ws.w_cs[k].as<int4>(), ctr[k].sched[0]);
        ktime_end(A, stream, kt, n_ptr);
        kt = ktime_begin(A, stream, KT_SHADE, n_k);
        k_shade<B><<<grid(g.shade, n_k), kBlockThreads, 0, stream>>>(
            S0, args, k, rays, n_ptr, ws.w_hits[k].as<WHit>(), ws.w_cs[k].as<int4>(), ws.w_nodes[k].as<WNode>(),
            ws.w_masks[k].as<unsigned>(), next, ctr + k, ws.w_shadow[k].as<unsigned>(), ws.w_slots[k].as<float4>());
        ktime_end(A, stream, kt, n_ptr);
        // The last level has no next level to overlap: its shadow rays stay
        // on the render stream (with that stream's spill area), where they
        // run beside the previous level's shadow rays on the side stream.
        const bool side_k = ov && k < pl.max_level, second = side_k && ov2 && (k & 1);
        hipStream_t sk = side_k ? (second ? so2 : so) : stream;
        if (side_k) {
            RTG_HIP(hipEventRecord(ws.ev_fork[k], stream));
            RTG_HIP(hipStreamWaitEvent(sk, ws.ev_fork[k], 0));
            (second ? forked2 : forked) = true;
        }
        if (pl.n_on > 0) {
            kt = ktime_begin(A, sk, KT_OCCLUDE, 0);
            k_occlude<B><<<grid(side_k ? g.occlude : g.occlude_alone, n_k * pl.n_on), kBlockThreads, 0, sk>>>(
                S0, side_k ? (second ? oargs2 : oargs) : args, ws.w_nodes[k].as<WNode>(),
                ws.w_shadow[k].as<unsigned>(), ctr + k, ws.w_masks[k].as<unsigned>());
            ktime_end(A, sk, kt, &ctr[k].n_shadow.v);
        }
        // Local lighting of every level but the last, once its shadow rays are
        // done, on their stream (beside the next levels): the level's
        // k_combine, on the frame's last stretch, then only composes.  Single
        // frames with every level enqueued at once only: where levels wait for
        // host read-backs (many lights, deep trees) the extra side work ended
        // up on the frame's path (measured).
        if (early_local && k < pl.max_level) {
            kt = ktime_begin(A, sk, KT_LOCAL, 0);
            k_local<B><<<lblocks, kBlockThreads, 0, sk>>>(S0, args, k, n_ptr, ws.w_nodes[k].as<WNode>(),
                                                          ws.w_masks[k].as<unsigned>());
            ktime_end(A, sk, kt, n_ptr);
        }
        RTG_HIP(hipGetLastError());
        last = k;
        if (k == pl.max_level) break;
        if (pl.no_sync) {
            n_k = pl.bound[k + 1];
            continue;
        }
        RTG_HIP(hipMemcpyAsync(ws.qcount_host, &ctr[k].n_next.v, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        RTG_HIP(hipStreamSynchronize(stream));
        n_k = ws.qcount_host[0];
    }
    {
        const int st = join();
        if (st != RTG_OK) return st;
    }
    if (cancel_host && *cancel_host) return RTG_OK;
    if (!pl.no_sync && ws.grow(ws.w_samples, (size_t)pl.n_samples * sizeof(float4)))
        return fail(RTG_E_NOMEM, "samples");
    const int blocks = std::max(8, cus * 16);
    for (int k = last; k >= 0; --k) {
        const WRay *rays = k == 0 ? nullptr : ws.w_rays[k].as<WRay>();
        const unsigned *n_ptr = k == 0 ? nullptr : &ctr[k - 1].n_next.v;
        const int kt = ktime_begin(A, stream, KT_COMBINE, k == 0 ? (uint64_t)pl.n_samples : 0);
        k_combine<B><<<blocks, kBlockThreads, 0, stream>>>(
            S0, args, k, rays, n_ptr, ws.w_nodes[k].as<WNode>(), ws.w_masks[k].as<unsigned>(),
            ws.w_slots[k].as<float4>(), k > 0 ? ws.w_slots[k - 1].as<float4>() : nullptr,
            ws.w_samples.as<float4>(), early_local && k < pl.max_level ? 1 : 0);
        ktime_end(A, stream, kt, n_ptr);
    }
    if (B || a.r0.spp != 1) {  // a single frame at one sample per pixel is written by k_combine
        const int kt = ktime_begin(A, stream, KT_RESOLVE, 0);
        k_resolve<B><<<dim3(B ? 256 : 1024, pl.n_reqs), 256, 0, stream>>>(args, ws.w_samples.as<float4>());
        ktime_end(A, stream, kt);
    }
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

static int host_buffers(Workspace &ws) {
    if (ws.qcount_host == nullptr)
        RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&ws.qcount_host), sizeof(unsigned) * 4, hipHostMallocDefault));
    return RTG_OK;
}

int reserve_frame(Workspace &ws, int W, int H, int spp, int max_level, int fanout, int n_on, Arena &A) {
    int st = host_buffers(ws);
    if (st == RTG_OK) st = ws.side_init();
    if (st != RTG_OK) return st;
    if (ws.grow(ws.tables, ((size_t)W + H) * sizeof(float))) return fail(RTG_E_NOMEM, "tables");
    if (ws.grow(ws.spill2, spill_ints(Grids<false>(A.cus > 0 ? A.cus : 256).occlude) * sizeof(int)))
        return fail(RTG_E_NOMEM, "spill");
    const size_t samples = frame_samples(W, H, spp, A.tune.tile_lg);
    // what launch_wavefront reserves up front: a no-sync plan of one pass
    if (fanout >= 2 || samples > (size_t)kNoSyncMaxSamples) return RTG_OK;
    const WavePlan pl = make_plan((int)samples, max_level, fanout, n_on, 1, A.no_sync_budget);
    return pl.no_sync ? reserve_all(ws, pl, A.cus > 0 ? A.cus : 256) : RTG_OK;
}

int launch_wavefront(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                     Workspace &ws, Arena &A, const Cancel *cancel) {
    const int *cancel_host = cancel ? cancel->host : nullptr;
    double dp;
    int st = upload_raygen_tables(S, W, H, stream, ws, dp);
    if (st != RTG_OK) return st;
    if ((st = host_buffers(ws)) != RTG_OK) return st;
    WaveArgs a{};
    WaveReq &r = a.r0;
    r.W = W, r.H = H, r.spp = p.spp > 0 ? p.spp : 1, r.max_level = p.max_level, r.tile_compat = p.tile_compat;
    const int lg = A.tune.tile_lg, tw = 1 << lg, th = 64 >> lg;
    r.tile_lg = lg;
    r.bw = (W + tw - 1) / tw;
    r.sample_base = 0;
    r.scene = 0;
    r.dpf = (float)dp;
    r.seed = p.seed;
    r.retest = (p.flags & RTG_FLAG_FP64_RETEST) != 0;
    r.no_refract = p.mode == RTG_MODE_REFLECTTEST;
    r.ax = ws.tables.as<float>();
    r.by = ws.tables.as<float>() + W;
    r.rgb = o.rgb, r.ids = o.ids, r.tprim = o.tprim;
    a.counts = o.counts;
    a.retests = o.retests;
    a.grab_lg = A.tune.grab_lg;
    // Frames whose ray trees can branch (refraction: two children per hit)
    // run in passes of whole block rows, at most kPassSamples level-0
    // samples each, so that the queues of all levels stay a few GB.
    const int rows = (H + th - 1) / th, per_row = r.bw * 64 * r.spp;
    const int n_on = __builtin_popcount(S.lights_on);
    int rows_per_pass = frame_rows_per_pass(rows, (size_t)per_row, S.max_fanout, n_on, kPassSamples);
    if (cancel && cancel->pass_samples > 0)
        rows_per_pass = std::min(rows_per_pass, std::max(1, (int)(cancel->pass_samples / (size_t)per_row)));
    for (int row0 = 0; row0 < rows; row0 += rows_per_pass) {
        r.row0 = row0;
        r.n_samples = std::min(rows_per_pass, rows - row0) * per_row;
        a.n_samples = r.n_samples;
        WavePlan pl = make_plan(r.n_samples, p.max_level, S.max_fanout, n_on, 1, A.no_sync_budget);
        // A very large frame keeps the GPU busy at every level; there the
        // host round trip per level costs little, and sizing each level from
        // the previous level's count saves the worst-case queue memory.
        if (r.n_samples > kNoSyncMaxSamples) pl.no_sync = false;
        if (pl.no_sync && (st = reserve_all(ws, pl, A.cus > 0 ? A.cus : 256)) != RTG_OK) return st;
        st = enqueue_wave<false>(S.view, a, pl, stream, ws, A, cancel, r.retest != 0,
                                 (p.flags & RTG_FLAG_ONE_STREAM) == 0);
        if (st != RTG_OK || (cancel_host && *cancel_host)) return st;
    }
    return RTG_OK;
}

// Pinned staging for batch tables; replaced only while `stream` is idle.
static int staging(Workspace &ws, size_t bytes, hipStream_t stream) {
    if (ws.staging_bytes >= bytes) return RTG_OK;
    RTG_HIP(hipStreamSynchronize(stream));
    if (ws.staging) (void)hipHostFree(ws.staging);
    ws.staging = nullptr;
    ws.staging_bytes = 0;
    const size_t n = std::max<size_t>(bytes, 1 << 16) * 5 / 4;
    RTG_HIP(hipHostMalloc(&ws.staging, n, hipHostMallocDefault));
    ws.staging_bytes = n;
    return RTG_OK;
}

// Pinned bytes a batch stages: ray-generation tables, request table, scene table.
static size_t batch_staging_bytes(size_t table_floats, size_t n_reqs, size_t n_scenes, size_t *off_reqs,
                                  size_t *off_scenes) {
    *off_reqs = (table_floats * sizeof(float) + 255) & ~size_t(255);
    *off_scenes = (*off_reqs + n_reqs * sizeof(WaveReq) + 255) & ~size_t(255);
    return *off_scenes + n_scenes * sizeof(SceneView);
}

int reserve_batch(Workspace &ws, size_t samples, int n_reqs, int n_scenes, size_t table_floats, int max_level,
                  int fanout, int n_on, hipStream_t stream, Arena &A) {
    int st = host_buffers(ws);
    if (st == RTG_OK) st = ws.side_init();
    if (st != RTG_OK) return st;
    if (samples == 0 || n_reqs <= 0) return RTG_OK;
    const size_t padded = batch_bucket(samples);
    if (padded > level_ray_cap(n_on) || n_reqs > kMaxBatch) return fail(RTG_E_LIMIT, "batch bound too large");
    int n_pad = 1;
    while (n_pad < n_reqs) n_pad <<= 1;
    const WavePlan pl = make_plan((int)padded, max_level, fanout, n_on, n_pad, A.no_sync_budget);
    if (pl.no_sync && (st = reserve_all(ws, pl, A.cus > 0 ? A.cus : 256)) != RTG_OK) return st;
    if (ws.grow(ws.b_reqs, (size_t)n_pad * sizeof(WaveReq)) || ws.grow(ws.b_scenes, (size_t)n_scenes * sizeof(SceneView)) ||
        ws.grow(ws.b_chunks, (padded >> 6) * sizeof(int)) || ws.grow(ws.tables, table_floats * sizeof(float)))
        return fail(RTG_E_NOMEM, "batch tables");
    size_t off_reqs, off_scenes;
    return staging(ws, batch_staging_bytes(table_floats, n_pad, n_scenes, &off_reqs, &off_scenes), stream);
}

int launch_batch(const BatchItem *items, int n, unsigned long long *counts, unsigned long long *retests,
                 hipStream_t stream, Workspace &ws, Arena &A, int capture_at, int *graph_replayed) {
    if (graph_replayed) *graph_replayed = 0;
    if (n <= 0) return RTG_OK;
    if (n > kMaxBatch) return fail(RTG_E_LIMIT, "batch larger than 4096 requests");
    // distinct scenes, request table, ray-generation tables
    std::vector<const SceneRes *> scenes;
    std::vector<int> scene_of(n);
    int fanout = 0, n_on = 0, max_level = 0;
    size_t table_floats = 0;
    for (int i = 0; i < n; ++i) {
        const SceneRes *s = items[i].scene;
        auto it = std::find(scenes.begin(), scenes.end(), s);
        scene_of[i] = (int)(it - scenes.begin());
        if (it == scenes.end()) scenes.push_back(s);
        fanout = std::max(fanout, s->max_fanout);
        n_on = std::max(n_on, __builtin_popcount(s->lights_on));
        max_level = std::max(max_level, (int)items[i].params.max_level);
        table_floats += (size_t)items[i].W + items[i].H;
    }
    // Pad the request count and the level-0 length to buckets so that
    // batches of similar size share one captured graph.
    int n_pad = 1;
    while (n_pad < n) n_pad <<= 1;
    std::vector<WaveReq> reqs(n_pad);
    size_t total = 0;
    bool any_retest = false, one_stream = false;
    for (int i = 0; i < n; ++i) {
        const BatchItem &it = items[i];
        WaveReq &r = reqs[i];
        r.W = it.W, r.H = it.H, r.spp = it.params.spp > 0 ? it.params.spp : 1;
        r.max_level = it.params.max_level, r.tile_compat = it.params.tile_compat;
        const int lg = A.tune.tile_lg;
        r.tile_lg = lg;
        r.bw = (it.W + (1 << lg) - 1) >> lg;
        r.n_samples = (int)frame_samples(it.W, it.H, r.spp, lg);
        r.sample_base = (int)total;
        r.scene = scene_of[i];
        r.seed = it.params.seed;
        r.retest = (it.params.flags & RTG_FLAG_FP64_RETEST) != 0;
        r.no_refract = it.params.mode == RTG_MODE_REFLECTTEST;
        any_retest = any_retest || r.retest;
        one_stream = one_stream || (it.params.flags & RTG_FLAG_ONE_STREAM) != 0;
        r.rgb = it.rgb, r.ids = it.ids, r.tprim = it.tprim;
        total += (size_t)r.n_samples;
    }
    const size_t padded = batch_bucket(total);
    if (padded > level_ray_cap(n_on))
        return fail(RTG_E_LIMIT, "batch has more level-0 samples than one wavefront level holds (" +
                                     std::to_string(padded) + " > " + std::to_string(level_ray_cap(n_on)) +
                                     "): split it");
    // Level 0 is read in whole 64-sample chunks (chunk_req holds one entry per
    // chunk): a length that is not a multiple of 64 would read past it.
    if (padded % 64 != 0 || padded < total) return fail(RTG_E_STATE, "batch bucket is not whole chunks");
    for (int i = n; i < n_pad; ++i) {
        WaveReq &r = reqs[i];
        r = WaveReq{};
        r.spp = 1;
        r.sample_base = (int)padded;
    }
    WavePlan pl = make_plan((int)padded, max_level, fanout, n_on, n_pad, A.no_sync_budget);
    bool use_graph = pl.no_sync && capture_at > 0;
    int st = host_buffers(ws);
    if (st != RTG_OK) return st;
    if (pl.no_sync && (st = reserve_all(ws, pl, A.cus > 0 ? A.cus : 256)) != RTG_OK) return st;
    const size_t n_chunks = padded >> 6;
    if (ws.grow(ws.b_reqs, reqs.size() * sizeof(WaveReq)) || ws.grow(ws.b_scenes, scenes.size() * sizeof(SceneView)) ||
        ws.grow(ws.b_chunks, n_chunks * sizeof(int)) || ws.grow(ws.tables, table_floats * sizeof(float)))
        return fail(RTG_E_NOMEM, "batch tables");
    ws.tables_key = 0;  // the single-frame table cache no longer holds
    // stage: ray-generation tables, request table, scene table
    size_t off_reqs, off_scenes;
    const size_t bytes = batch_staging_bytes(table_floats, reqs.size(), scenes.size(), &off_reqs, &off_scenes);
    if ((st = staging(ws, bytes, stream)) != RTG_OK) return st;
    RTG_HIP(hipStreamSynchronize(stream));  // the staging buffer may still feed the previous batch's copies
    uint8_t *stg = static_cast<uint8_t *>(ws.staging);
    float *tab = reinterpret_cast<float *>(stg);
    size_t toff = 0;
    for (int i = 0; i < n; ++i) {
        const double dp = raygen_tables(items[i].scene->camera.fovy, items[i].W, items[i].H, tab + toff);
        reqs[i].dpf = (float)dp;
        reqs[i].ax = ws.tables.as<float>() + toff;
        reqs[i].by = ws.tables.as<float>() + toff + items[i].W;
        toff += (size_t)items[i].W + items[i].H;
    }
    std::memcpy(stg + off_reqs, reqs.data(), reqs.size() * sizeof(WaveReq));
    for (size_t s = 0; s < scenes.size(); ++s)
        std::memcpy(stg + off_scenes + s * sizeof(SceneView), &scenes[s]->view, sizeof(SceneView));
    RTG_HIP(hipMemcpyAsync(ws.tables.p, stg, table_floats * sizeof(float), hipMemcpyHostToDevice, stream));
    RTG_HIP(hipMemcpyAsync(ws.b_reqs.p, stg + off_reqs, reqs.size() * sizeof(WaveReq), hipMemcpyHostToDevice, stream));
    RTG_HIP(hipMemcpyAsync(ws.b_scenes.p, stg + off_scenes, scenes.size() * sizeof(SceneView), hipMemcpyHostToDevice,
                           stream));
    WaveArgs a{};
    a.reqs = ws.b_reqs.as<WaveReq>();
    a.scenes = ws.b_scenes.as<SceneView>();
    a.chunk_req = ws.b_chunks.as<int>();
    a.n_samples = (int)padded;
    a.grab_lg = A.tune.grab_lg;
    a.counts = counts;
    a.retests = retests;
    const uint64_t key = (uint64_t)padded ^ ((uint64_t)n_pad << 32) ^ ((uint64_t)max_level << 45) ^
                         ((uint64_t)fanout << 50) ^ ((uint64_t)n_on << 53) ^ ((uint64_t)any_retest << 60) ^
                         (ws.epoch * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uintptr_t)counts * 0xC2B2AE3D27D4EB4Full) ^
                         ((uint64_t)(uintptr_t)retests * 0x165667B19E3779F9ull) ^
                         (A.tune.key() * 0x27D4EB2F165667C5ull) ^ ((uint64_t)one_stream << 63);
    // each level's shadow rays beside the next level on the side stream
    // (`overlap`), or all on `stream`
    auto enqueue = [&](bool overlap) -> int {
        k_chunk_req<<<(int)((n_chunks + 255) / 256), 256, 0, stream>>>(a.reqs, n_pad, ws.b_chunks.as<int>(),
                                                                      (int)n_chunks);
        return enqueue_wave<true>(SceneView{}, a, pl, stream, ws, A, nullptr, any_retest, overlap && !one_stream);
    };
    if (use_graph && capture_at > 1 && !(ws.graph != nullptr && ws.graph_key == key)) {
        int slot = -1;
        for (int i = 0; i < Workspace::kSeen; ++i)
            if (ws.seen_count[i] && ws.seen_keys[i] == key) slot = i;
        if (slot < 0) {
            slot = ws.seen_next++ % Workspace::kSeen;
            ws.seen_keys[slot] = key;
            ws.seen_count[slot] = 0;
        }
        if ((int)++ws.seen_count[slot] < capture_at) use_graph = false;
    }
    if (!use_graph) return enqueue(true);
    if (ws.graph == nullptr || ws.graph_key != key) {
        if (ws.graph) (void)hipGraphExecDestroy(ws.graph);
        ws.graph = nullptr;
        const bool kt = A.ktimes;
        A.ktimes = false;  // no timing events inside a captured graph
        // Capture with the side stream (the fork and join events become graph
        // edges); should that capture fail, capture the one-stream order.
        std::string why;
        for (int attempt = 0; attempt < 2 && ws.graph == nullptr; ++attempt) {
            RTG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            st = enqueue(attempt == 0);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            if (st != RTG_OK) {
                if (g) (void)hipGraphDestroy(g);
                A.ktimes = kt;
                return st;
            }
            if (e != hipSuccess) {
                why = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
                (void)hipGetLastError();
                continue;
            }
            const hipError_t ei = hipGraphInstantiate(&ws.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                ws.graph = nullptr;
                why = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
                (void)hipGetLastError();
            }
        }
        A.ktimes = kt;
        if (ws.graph == nullptr) return fail(RTG_E_DEVICE, why);
        ws.graph_key = key;
    } else if (graph_replayed) {
        *graph_replayed = 1;
    }
    RTG_HIP(hipGraphLaunch(ws.graph, stream));
    return RTG_OK;
}

}  // namespace rtg
