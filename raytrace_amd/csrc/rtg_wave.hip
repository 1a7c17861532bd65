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
//   resolve average of the samples of a pixel, quantised to RGB8
// The per-node arithmetic is the megakernel's (shade_local, compose), so the
// two engines give bit-identical frames.
#include "rt_devstack.hpp"
#include "rtg_internal.hpp"

namespace rtg {

struct WRay {  // 48 B
    float ox, oy, oz, dx, dy, dz, bwc, mtlrfr;
    int link;  // 2 * parent node + slot (0 reflection, 1 refraction)
    int src_obj, src_prim;
    int flags;  // bits 0-2 type, bit 3 inside, bits 4-6 source copy bin
};
struct WHit {  // 32 B
    float t, w, u, v;
    int kind, obj, prim, pad;
};
struct WNode {  // 64 B
    float px, py, pz, nx, ny, nz, dx, dy, dz, vr, vg, vb, t;
    int mtl, id_prim;
    int bits;  // bits 0-23 object, 24-26 copy bin, 27 valid, 28 inside, 29 total internal reflection
};

// Device counters of one level, zeroed at the start of a frame; each in its
// own 256-byte line.
struct LevelCtr {
    LineCtr n_next;       // child rays appended for level k + 1
    LineCtr n_shadow;     // shadow-ray entries appended
    LineCtr sched[3][8];  // WaveQueue counters of extend, shade, occlude
};
constexpr int kCtrWords = (int)(sizeof(LevelCtr) / sizeof(unsigned));

struct WaveArgs {
    int W, H, spp, max_level, tile_compat;
    int bw;         // 8x8 pixel blocks per row
    int n_samples;  // pixel blocks * 64 * spp
    float dpf;
    uint64_t seed;
    const float *ax, *by;
    int32_t *ids;
    float *tprim;
    unsigned long long *counts;
    int *spill;
};

__device__ inline bool sample_pixel(const WaveArgs &A, int s, int &px, int &py, int &si) {
    const int per_block = 64 * A.spp;
    const int pb = s / per_block, r = s - pb * per_block;
    const int lane = r / A.spp;
    si = r - lane * A.spp;
    px = (pb % A.bw) * 8 + (lane & 7);
    py = (pb / A.bw) * 8 + (lane >> 3);
    if (px >= A.W || py >= A.H) return false;
    if (A.tile_compat && (px >= ((A.W >> 6) << 6) || py >= ((A.H >> 6) << 6))) return false;
    return true;
}

__device__ inline TraceRay primary_ray_w(const SceneView &S, const WaveArgs &A, int px, int py, int s) {
    float a = A.ax[px], b = A.by[py];
    if (A.spp > 1) {
        const uint64_t h = mix64(A.seed ^ ((uint64_t)(py * A.W + px) * 0x9E3779B97F4A7C15ull) ^
                                 ((uint64_t)s * 0xC2B2AE3D27D4EB4Full));
        const float jx = (float)(h >> 40) * (1.0f / 16777216.0f);
        const float jy = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        a = a + (((float)(s & 1) + jx) * 0.5f) * A.dpf;
        b = b + (((float)((s >> 1) & 1) + jy) * 0.5f) * A.dpf;
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

__device__ inline WRay make_wray(const TraceRay &q, float bwc, float mtlrfr, int link) {
    WRay w;
    w.ox = q.r.o.x, w.oy = q.r.o.y, w.oz = q.r.o.z;
    w.dx = q.r.d.x, w.dy = q.r.d.y, w.dz = q.r.d.z;
    w.bwc = bwc, w.mtlrfr = mtlrfr;
    w.link = link;
    w.src_obj = q.src.obj, w.src_prim = q.src.prim;
    w.flags = (q.type & 7) | ((q.inside & 1) << 3) | ((q.src.bin & 7) << 4);
    return w;
}

__device__ inline void add_counts(unsigned long long *counts, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    a = wave_sum_u32(a), b = wave_sum_u32(b), c = wave_sum_u32(c), d = wave_sum_u32(d);
    if (__lane_id() == 0) {
        if (a) atomicAdd(&counts[0], (unsigned long long)a);
        if (b) atomicAdd(&counts[1], (unsigned long long)b);
        if (c) atomicAdd(&counts[2], (unsigned long long)c);
        if (d) atomicAdd(&counts[3], (unsigned long long)d);
    }
}

__global__ __launch_bounds__(kBlockThreads) void k_extend(SceneView S, WaveArgs A, int k, const WRay *rays,
                                                          const unsigned *n_ptr, WHit *hits, LineCtr *sched) {
    __shared__ int lds[kLdsStack * kBlockThreads];
    DevStack st = make_stack(lds, A.spill);
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    WaveQueue wq(sched, n);
    uint32_t c_prim = 0, c_refl = 0, c_refr = 0;
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int i = base + __lane_id();
        if (i >= n) continue;
        TraceRay q;
        int px = 0, py = 0, si = 0;
        bool valid = true;
        if (k == 0) {
            valid = sample_pixel(A, i, px, py, si);
            if (valid) q = primary_ray_w(S, A, px, py, si);
        } else {
            float bwc, mtlrfr;
            q = load_ray(rays[i], bwc, mtlrfr);
            if (q.src.obj >= 0) q.self_test = self_testable(S, q.r, q.src, kMiss);
        }
        Hit h{kMiss, HIT_NONE, -1, -1, 0, 0, 0};
        if (valid) {
            h = trace_closest(S, q, st);
            if (k == 0)
                ++c_prim;
            else if (q.type == RAY_REFLECT)
                ++c_refl;
            else
                ++c_refr;
        }
        hits[i] = WHit{h.t, h.w, h.u, h.v, h.kind, h.obj, h.prim, 0};
        if (k == 0 && valid && si == 0) {
            const size_t pix = (size_t)py * A.W + px;
            if (A.ids) {
                A.ids[2 * pix] = h.obj;
                A.ids[2 * pix + 1] = id_prim(S, h);
            }
            if (A.tprim) A.tprim[pix] = h.t;
        }
    }
    add_counts(A.counts, c_prim, 0, c_refl, c_refr);
}

__global__ __launch_bounds__(kBlockThreads) void k_shade(SceneView S, WaveArgs A, int k, const WRay *rays,
                                                         const unsigned *n_ptr, const WHit *hits, WNode *nodes,
                                                         unsigned *masks, WRay *next, LevelCtr *ctr, unsigned *shadow,
                                                         uint32_t lights_on) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    const int n_on = __builtin_popcount(lights_on);
    WaveQueue wq(ctr->sched[1], n);
    uint32_t c_shadow = 0;
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int i = base + __lane_id();
        int n_child = 0;
        bool lit = false;
        WRay c0, c1;  // children in append order (registers, no indexed array)
        V3 P = V3{0, 0, 0}, N = V3{0, 0, 0}, D = V3{0, 0, 0};
        if (i < n) {
            TraceRay q;
            float bwc = 1.0f, mtlrfr = 1.0f;
            bool valid = true;
            if (k == 0) {
                int px, py, si;
                valid = sample_pixel(A, i, px, py, si);
                if (valid) q = primary_ray_w(S, A, px, py, si);
            } else {
                q = load_ray(rays[i], bwc, mtlrfr);
            }
            const WHit wh = hits[i];
            const float zNear = k == 0 ? S.cam.zNear : 0.0f;
            WNode nd;
            nd.bits = 0;
            if (valid && wh.kind != HIT_NONE && !(wh.t > S.cam.zFar || wh.t < zNear)) {
                const Hit h{wh.t, wh.kind, wh.obj, wh.prim, wh.w, wh.u, wh.v};
                const HitRec hr = make_hitrec(S, q, h);
                const V3 vc = texel(S, hr.tex, hr.tu, hr.tv);
                const DevMaterial &m = S.mtls[hr.mtl];
                int tir = 0;
                if (m.reflect > 0.01f && k + 1 <= A.max_level && bwc * m.reflect >= 1e-5f) {
                    c0 = make_wray(reflect_ray(q, hr), bwc * m.reflect, 1.0f, 2 * i);
                    n_child = 1;
                }
                if (m.refract > 0.01f) {
                    TraceRay cq;
                    if (!refract_ray(q, mtlrfr, hr, cq)) {
                        tir = 1;
                    } else if (k + 1 <= A.max_level && bwc * m.refract >= 1e-5f) {
                        const WRay w = make_wray(cq, bwc * m.refract, hr.rfr, 2 * i + 1);
                        if (n_child == 0)
                            c0 = w;
                        else
                            c1 = w;
                        ++n_child;
                    }
                }
                nd.px = hr.P.x, nd.py = hr.P.y, nd.pz = hr.P.z;
                nd.nx = hr.N.x, nd.ny = hr.N.y, nd.nz = hr.N.z;
                nd.dx = q.r.d.x, nd.dy = q.r.d.y, nd.dz = q.r.d.z;
                nd.vr = vc.x, nd.vg = vc.y, nd.vb = vc.z;
                nd.t = hr.t;
                nd.mtl = hr.mtl;
                nd.id_prim = hr.id.prim;
                nd.bits = (hr.id.obj & 0xFFFFFF) | ((hr.id.bin & 7) << 24) | (1 << 27) | ((hr.inside & 1) << 28) |
                          (tir << 29);
                lit = true;
                P = hr.P, N = hr.N, D = q.r.d;
                c_shadow += n_on;  // the reference traces one shadow ray per switched-on light
            }
            nodes[i] = nd;
            masks[i] = 0;
        }
        const int cb = wave_append(&ctr->n_next.v, n_child);
        if (n_child > 0) next[cb] = c0;
        if (n_child > 1) next[cb + 1] = c1;
        // Shadow entries (node << 5 | light), light-major within the wave's
        // block, so that a wave of k_occlude traces neighbouring points toward
        // one light.  Lights that cannot change the colour are not traced
        // (shade_local ignores their mask bit).
        uint32_t emit = 0;
        if (lit) {
            int r = 0;
            for (uint32_t lm = lights_on; lm; lm &= lm - 1, ++r)
                if (light_dots(N, D, light_term(S.lights[__builtin_ctz(lm)], P)).matters()) emit |= 1u << r;
        }
        int total = 0;
        for (int r = 0; r < n_on; ++r) total += __popcll(__ballot((emit >> r) & 1u));
        int sb = 0;
        if (__lane_id() == 0 && total > 0) sb = (int)atomicAdd(&ctr->n_shadow.v, (unsigned)total);
        sb = __shfl(sb, 0, 64);
        const unsigned long long below = (1ull << __lane_id()) - 1ull;
        int r = 0;
        for (uint32_t lm = lights_on; lm; lm &= lm - 1, ++r) {
            const unsigned long long m = __ballot((emit >> r) & 1u);
            if ((emit >> r) & 1u) shadow[sb + __popcll(m & below)] = ((unsigned)i << 5) | (unsigned)__builtin_ctz(lm);
            sb += __popcll(m);
        }
    }
    add_counts(A.counts, 0, c_shadow, 0, 0);
}

__device__ inline SrcId node_id(const WNode &nd) { return SrcId{nd.bits & 0xFFFFFF, nd.id_prim, (nd.bits >> 24) & 7}; }

__global__ __launch_bounds__(kBlockThreads) void k_occlude(SceneView S, WaveArgs A, const WNode *nodes,
                                                           const unsigned *shadow, LevelCtr *ctr, unsigned *masks) {
    __shared__ int lds[kLdsStack * kBlockThreads];
    DevStack st = make_stack(lds, A.spill);
    const int n = (int)ctr->n_shadow.v;
    WaveQueue wq(ctr->sched[2], n);
    for (int base = wq.next(); base >= 0; base = wq.next()) {
        const int j = base + __lane_id();
        if (j >= n) continue;
        const unsigned e = shadow[j];
        const int i = (int)(e >> 5), li = (int)(e & 31);
        const WNode nd = nodes[i];
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

__global__ void k_fill_slots(float4 *slots, const unsigned *n_ptr, int n_fixed) {
    const int n = 2 * (n_fixed >= 0 ? n_fixed : (int)*n_ptr);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        slots[i] = make_float4(0.0f, 0.0f, 0.0f, kMiss);
}

__global__ __launch_bounds__(kBlockThreads) void k_combine(SceneView S, WaveArgs A, int k, const WRay *rays,
                                                           const unsigned *n_ptr, const WNode *nodes,
                                                           const unsigned *masks, const float4 *slots,
                                                           float4 *parent_slots, float4 *samples) {
    const int n = k == 0 ? A.n_samples : (int)*n_ptr;
    const ChunkRange cr = xcd_chunks((n + kBlockThreads - 1) / kBlockThreads);
    for (int c = cr.first; c < cr.end; c += cr.step) {
        const int i = c * kBlockThreads + threadIdx.x;
        if (i >= n) continue;
        const WNode nd = nodes[i];
        float4 out = make_float4(0.0f, 0.0f, 0.0f, kMiss);
        if ((nd.bits >> 27) & 1) {
            const unsigned mask = masks[i];
            const V3 L = shade_local(S, V3{nd.dx, nd.dy, nd.dz}, V3{nd.px, nd.py, nd.pz}, V3{nd.nx, nd.ny, nd.nz},
                                     nd.mtl, V3{nd.vr, nd.vg, nd.vb},
                                     [&](int li, const LightTerm &, bool) -> bool { return (mask >> li) & 1u; });
            const DevMaterial &m = S.mtls[nd.mtl];
            const float4 a = slots[2 * i], b = slots[2 * i + 1];
            const V3 col = compose(L, m.reflect, m.refract, (nd.bits >> 29) & 1, (nd.bits >> 28) & 1, m.kd,
                                   V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, b.w);
            out = make_float4(col.x, col.y, col.z, nd.t);
        }
        if (k == 0)
            samples[i] = out;
        else
            parent_slots[rays[i].link] = out;
    }
}

__global__ void k_resolve(WaveArgs A, const float4 *samples, uint8_t *rgb) {
    const int npix = A.W * A.H;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        const int px = p % A.W, py = p / A.W;
        if (A.tile_compat && (px >= ((A.W >> 6) << 6) || py >= ((A.H >> 6) << 6))) continue;
        const int pb = (py >> 3) * A.bw + (px >> 3), lane = ((py & 7) << 3) | (px & 7);
        const int s0 = (pb * 64 + lane) * A.spp;
        V3 acc = V3{0, 0, 0};
        for (int s = 0; s < A.spp; ++s) {
            const float4 c = samples[s0 + s];
            acc = acc + V3{c.x, c.y, c.z};
        }
        if (A.spp > 1) acc = acc * (1.0f / (float)A.spp);
        rgb[3 * (size_t)p + 0] = quantize(acc.x);
        rgb[3 * (size_t)p + 1] = quantize(acc.y);
        rgb[3 * (size_t)p + 2] = quantize(acc.z);
    }
}

__global__ void k_zero_u32(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Largest queue memory (bytes) a frame may reserve to run without host round
// trips between levels.
constexpr size_t kNoSyncBudget = size_t(8) << 30;

// Resident blocks per CU of a kernel at kBlockThreads threads (registers and
// LDS permitting), queried once.
template <class K>
static int blocks_per_cu(K kernel) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlockThreads, 0) != hipSuccess || nb < 1) nb = 4;
    return nb;
}

int launch_wavefront(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                     Arena &A) {
    double dp;
    int st = upload_raygen_tables(S, W, H, stream, A, dp);
    if (st != RTG_OK) return st;
    const int cus = A.cus > 0 ? A.cus : 256;
    static const int occ_extend = blocks_per_cu(k_extend), occ_shade = blocks_per_cu(k_shade),
                     occ_occlude = blocks_per_cu(k_occlude);
    const int g_extend = cus * occ_extend, g_shade = cus * occ_shade, g_occlude = cus * occ_occlude;
    if (ensure(A.spill, spill_ints(std::max(g_extend, g_occlude)) * sizeof(int))) return fail(RTG_E_NOMEM, "spill");
    constexpr size_t kCtrBytes = (size_t)Arena::kLevels * sizeof(LevelCtr);
    if (A.qcount_host == nullptr)
        RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&A.qcount_host), kCtrBytes, hipHostMallocDefault));
    if (ensure(A.w_qcount, kCtrBytes)) return fail(RTG_E_NOMEM, "queue counters");
    WaveArgs a;
    a.W = W, a.H = H, a.spp = p.spp > 0 ? p.spp : 1, a.max_level = p.max_level, a.tile_compat = p.tile_compat;
    a.bw = (W + 7) / 8;
    a.n_samples = a.bw * ((H + 7) / 8) * 64 * a.spp;
    a.dpf = (float)dp;
    a.seed = p.seed;
    a.ax = A.tables.as<float>();
    a.by = A.tables.as<float>() + W;
    a.ids = o.ids, a.tprim = o.tprim;
    a.counts = o.counts;
    a.spill = A.spill.as<int>();
    const uint32_t lights_on = S.lights_on;  // lights that trace shadow rays (Light::bLight)
    const int n_on = __builtin_popcount(lights_on);
    LevelCtr *ctr = A.w_qcount.as<LevelCtr>();
    const int ctr_words = Arena::kLevels * kCtrWords;
    k_zero_u32<<<(ctr_words + 255) / 256, 256, 0, stream>>>(A.w_qcount.as<unsigned>(), ctr_words);
    auto grid = [](int g, size_t items) { return (int)std::max<size_t>(8, std::min<size_t>(g, (items + 127) / 128)); };
    // Queue capacity per level: level k holds at most n_samples * fanout^k
    // rays.  When those bounds fit the budget, every queue is sized by its
    // bound and the frame is enqueued without a host round trip (kernels read
    // the live counts from the device); otherwise each level is sized from
    // the previous level's count, read back to the host.
    const size_t per_ray = sizeof(WRay) + sizeof(WHit) + sizeof(WNode) + sizeof(unsigned) + 2 * sizeof(float4) +
                           (size_t)n_on * sizeof(unsigned);
    size_t bound[Arena::kLevels] = {};
    size_t need = 0;
    bool no_sync = true;
    for (int k = 0; k <= p.max_level; ++k) {
        bound[k] = k == 0 ? (size_t)a.n_samples : bound[k - 1] * (size_t)S.max_fanout;
        need += bound[k] * per_ray;
        if (need > kNoSyncBudget) no_sync = false;
    }
    size_t n_k = (size_t)a.n_samples;
    int last = -1;
    for (int k = 0; k <= p.max_level && n_k > 0; ++k) {
        if (*A.cancel_host) return RTG_OK;
        const size_t cap_next = k < p.max_level ? (size_t)std::max(1, S.max_fanout) * n_k : 1;
        if (ensure(A.w_hits[k], n_k * sizeof(WHit)) || ensure(A.w_nodes[k], n_k * sizeof(WNode)) ||
            ensure(A.w_masks[k], n_k * sizeof(unsigned)) || ensure(A.w_slots[k], 2 * n_k * sizeof(float4)) ||
            ensure(A.w_shadow[k], std::max<size_t>(1, n_k * n_on) * sizeof(unsigned)) ||
            ensure(A.w_rays[k + 1 < Arena::kLevels ? k + 1 : k], cap_next * sizeof(WRay)))
            return fail(RTG_E_NOMEM, "wavefront queues");
        const WRay *rays = k == 0 ? nullptr : A.w_rays[k].as<WRay>();
        const unsigned *n_ptr = k == 0 ? nullptr : &ctr[k - 1].n_next.v;
        WRay *next = A.w_rays[k + 1 < Arena::kLevels ? k + 1 : k].as<WRay>();
        int kt = ktime_begin(A, stream, KT_EXTEND, n_k);
        k_extend<<<grid(g_extend, n_k), kBlockThreads, 0, stream>>>(S.view, a, k, rays, n_ptr,
                                                                    A.w_hits[k].as<WHit>(), ctr[k].sched[0]);
        ktime_end(A, stream, kt);
        kt = ktime_begin(A, stream, KT_SHADE, n_k);
        k_shade<<<grid(g_shade, n_k), kBlockThreads, 0, stream>>>(
            S.view, a, k, rays, n_ptr, A.w_hits[k].as<WHit>(), A.w_nodes[k].as<WNode>(), A.w_masks[k].as<unsigned>(),
            next, ctr + k, A.w_shadow[k].as<unsigned>(), lights_on);
        ktime_end(A, stream, kt);
        k_fill_slots<<<256, 256, 0, stream>>>(A.w_slots[k].as<float4>(), n_ptr, k == 0 ? a.n_samples : -1);
        if (n_on > 0) {
            kt = ktime_begin(A, stream, KT_OCCLUDE, 0);
            k_occlude<<<grid(g_occlude, n_k * n_on), kBlockThreads, 0, stream>>>(
                S.view, a, A.w_nodes[k].as<WNode>(), A.w_shadow[k].as<unsigned>(), ctr + k,
                A.w_masks[k].as<unsigned>());
            ktime_end(A, stream, kt, &ctr[k].n_shadow.v);
        }
        RTG_HIP(hipGetLastError());
        last = k;
        if (k == p.max_level) break;
        if (no_sync) {
            n_k = bound[k + 1];
            continue;
        }
        RTG_HIP(hipMemcpyAsync(A.qcount_host, &ctr[k].n_next.v, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        RTG_HIP(hipStreamSynchronize(stream));
        n_k = A.qcount_host[0];
    }
    if (*A.cancel_host) return RTG_OK;
    if (ensure(A.w_samples, (size_t)a.n_samples * sizeof(float4))) return fail(RTG_E_NOMEM, "samples");
    const int blocks = std::max(8, cus * 8);
    for (int k = last; k >= 0; --k) {
        const WRay *rays = k == 0 ? nullptr : A.w_rays[k].as<WRay>();
        const unsigned *n_ptr = k == 0 ? nullptr : &ctr[k - 1].n_next.v;
        const int kt = ktime_begin(A, stream, KT_COMBINE, k == 0 ? (uint64_t)a.n_samples : 0);
        k_combine<<<blocks, kBlockThreads, 0, stream>>>(
            S.view, a, k, rays, n_ptr, A.w_nodes[k].as<WNode>(), A.w_masks[k].as<unsigned>(),
            A.w_slots[k].as<float4>(), k > 0 ? A.w_slots[k - 1].as<float4>() : nullptr, A.w_samples.as<float4>());
        ktime_end(A, stream, kt, n_ptr);
    }
    const int kt = ktime_begin(A, stream, KT_RESOLVE, (uint64_t)W * H);
    k_resolve<<<1024, 256, 0, stream>>>(a, A.w_samples.as<float4>(), o.rgb);
    ktime_end(A, stream, kt);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

}  // namespace rtg
