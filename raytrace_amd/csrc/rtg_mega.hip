// rtg_mega.hip — the one-launch-per-frame megakernel.
//
// Persistent grid (8 blocks per CU), 128 threads per block = two waves, each
// wave an 8x8 pixel block; pixel blocks are split over the XCDs by
// xcd_chunks so each XCD's L2 caches the part of the BVH its image region
// touches.  Each thread evaluates its pixel's whole ray tree (render_sample).
#include "rt_devstack.hpp"
#include "rtg_internal.hpp"

namespace rtg {

struct MegaArgs {
    int W, H, spp, max_level, tile_compat;
    int mode;  // RTG_MODE_*: the full ray tree, the reflect test, or a debug mode
    int bw, nblocks;
    float dpf;
    uint64_t seed;
    const float *ax, *by;
    uint8_t *rgb;
    int32_t *ids;
    float *tprim;
    unsigned long long *counts;
    const int *cancel;
    int *spill;
};

// Primary ray of sample s of pixel (px, py): the reference's pixel-corner ray
// for spp 1 (RayTracer.cpp:26-27), 2x2 jittered strata otherwise.
__device__ inline TraceRay primary_ray(const SceneView &S, const float *ax, const float *by, int W, int px, int py,
                                       int s, int spp, float dpf, uint64_t seed) {
    float a = ax[px], b = by[py];
    if (spp > 1) {
        const uint64_t h =
            mix64(seed ^ ((uint64_t)(py * W + px) * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)s * 0xC2B2AE3D27D4EB4Full));
        const float jx = (float)(h >> 40) * (1.0f / 16777216.0f);
        const float jy = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        a = a + (((float)(s & 1) + jx) * 0.5f) * dpf;
        b = b + (((float)((s >> 1) & 1) + jy) * 0.5f) * dpf;
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

__global__ __launch_bounds__(kBlockThreads) void k_megakernel(SceneView S, MegaArgs A) {
    __shared__ int lds[kLdsStack * kBlockThreads];
    __shared__ Bvh4Node top_nodes[kTopNodes];
    const int tid = threadIdx.x;
    DevStack st = make_stack(lds, A.spill);
    stack_top_nodes(st, top_nodes, S);
    RayCounts rc{0, 0, 0, 0, 0};
    const ChunkRange cr = xcd_chunks(A.nblocks);
    const int tw = (A.W >> 6) << 6, th = (A.H >> 6) << 6;
    for (int pb = cr.first; pb < cr.end; pb += cr.step) {
        if (A.cancel != nullptr && __builtin_nontemporal_load(A.cancel) != 0) break;
        const int bx = pb % A.bw, by = pb / A.bw;
        const int px = bx * 16 + ((tid >> 6) << 3) + (tid & 7), py = by * 8 + ((tid & 63) >> 3);
        if (px >= A.W || py >= A.H) continue;
        if (A.tile_compat && (px >= tw || py >= th)) continue;
        V3 acc = V3{0, 0, 0};
        PrimaryOut po{-1, -1, kMiss};
        if (A.mode == RTG_MODE_CHECK) {  // RTcheck (RayTracer.cpp:46-77): tile checkerboard, no rays
            const float v = ((py >> 6) & 1) == ((px >> 6) & 1) ? 1.0f : 0.0f;
            acc = V3{v, v, v};
        } else {
            const bool debug = A.mode >= RTG_MODE_DEPTHTEST && A.mode <= RTG_MODE_SHADOWTEST;
            for (int s = 0; s < A.spp; ++s) {
                const TraceRay q = primary_ray(S, A.ax, A.by, A.W, px, py, s, A.spp, A.dpf, A.seed);
                PrimaryOut p0;
                const V3 c = debug ? debug_sample(S, A.mode, q, st, rc, p0)
                                   : render_sample(S, A.max_level, q, st, rc, p0, A.mode == RTG_MODE_REFLECTTEST);
                if (s == 0) po = p0;
                acc = acc + c;
            }
        }
        if (A.spp > 1 && A.mode != RTG_MODE_CHECK) acc = acc * (1.0f / (float)A.spp);
        const size_t pix = (size_t)py * A.W + px;
        A.rgb[3 * pix + 0] = quantize(acc.x);
        A.rgb[3 * pix + 1] = quantize(acc.y);
        A.rgb[3 * pix + 2] = quantize(acc.z);
        if (A.ids) {
            A.ids[2 * pix] = po.obj;
            A.ids[2 * pix + 1] = po.prim;
        }
        if (A.tprim) A.tprim[pix] = po.t;
    }
    const uint32_t c0 = wave_sum_u32(rc.primary), c1 = wave_sum_u32(rc.shadow), c2 = wave_sum_u32(rc.reflect),
                   c3 = wave_sum_u32(rc.refract);
    if ((tid & 63) == 0) {
        atomicAdd(&A.counts[0], (unsigned long long)c0);
        atomicAdd(&A.counts[1], (unsigned long long)c1);
        atomicAdd(&A.counts[2], (unsigned long long)c2);
        atomicAdd(&A.counts[3], (unsigned long long)c3);
    }
}

// Ray-generation tables (float)(xcur * dp), (float)(ycur * dp) with
// dp = tan(fovy * PI / 360) / (H / 2) in double (RayTracer.cpp:14,26).
// Uploaded once per (W, H, fovy); later frames reuse them.
double raygen_tables(float fovy, int W, int H, float *tab) {
    const double dp = tan((double)fovy * 3.1415926535897932384 / 360) / (H / 2);
    for (int x = 0; x < W; ++x) tab[x] = (float)((x - W / 2) * dp);
    for (int y = 0; y < H; ++y) tab[W + y] = (float)((y - H / 2) * dp);
    return dp;
}

// The same tables written on the device: (x - W/2) * dp is one correctly
// rounded double product, rounded to float, as on the host.
__global__ void k_raygen_tables(float *tab, int W, int H, double dp) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < W + H; i += gridDim.x * blockDim.x)
        tab[i] = i < W ? (float)((double)(i - W / 2) * dp) : (float)((double)(i - W - H / 2) * dp);
}

int write_raygen_tables(float *tab, int W, int H, double dp, hipStream_t stream) {
    k_raygen_tables<<<(W + H + 255) / 256, 256, 0, stream>>>(tab, W, H, dp);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

int upload_raygen_tables(const SceneRes &S, int W, int H, hipStream_t stream, Workspace &ws, double &dp) {
    dp = tan((double)S.camera.fovy * 3.1415926535897932384 / 360) / (H / 2);
    const uint32_t fb = __builtin_bit_cast(uint32_t, S.camera.fovy);
    const uint64_t key = (uint64_t)W | ((uint64_t)H << 16) | ((uint64_t)fb << 32);
    if (ws.tables.p != nullptr && ws.tables_key == key) return RTG_OK;
    const size_t bytes = ((size_t)W + H) * sizeof(float);
    if (ws.tables.p == nullptr || ws.tables.bytes < bytes) {
        RTG_HIP(hipStreamSynchronize(stream));  // earlier frames may still read the old tables
        if (ws.grow(ws.tables, bytes)) return fail(RTG_E_NOMEM, "tables");
    }
    // ordered after the earlier frames on `stream`, which read the old tables
    k_raygen_tables<<<(W + H + 255) / 256, 256, 0, stream>>>(ws.tables.as<float>(), W, H, dp);
    RTG_HIP(hipGetLastError());
    ws.tables_key = key;
    return RTG_OK;
}

int persistent_blocks(const Arena &A) { return std::max(8, (A.cus > 0 ? A.cus : 256) * 8); }

int enqueue_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                       Arena &A, const Cancel *cancel, const MegaScratch &ms) {
    MegaArgs a;
    a.W = W, a.H = H, a.spp = p.spp > 0 ? p.spp : 1, a.max_level = p.max_level, a.tile_compat = p.tile_compat;
    a.mode = p.mode;
    a.bw = (W + 15) / 16;
    a.nblocks = a.bw * ((H + 7) / 8);
    a.dpf = (float)ms.dp;
    a.seed = p.seed;
    a.ax = ms.tables;
    a.by = ms.tables + W;
    a.rgb = o.rgb, a.ids = o.ids, a.tprim = o.tprim;
    a.counts = o.counts;
    a.cancel = cancel ? cancel->dev : nullptr;
    a.spill = ms.spill;
    const int kt = ktime_begin(A, stream, KT_MEGA, (uint64_t)W * H * a.spp);
    k_megakernel<<<ms.blocks, kBlockThreads, 0, stream>>>(S.view, a);
    ktime_end(A, stream, kt);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

int launch_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                      Arena &A, const Cancel *cancel) {
    MegaScratch ms{};
    int st = upload_raygen_tables(S, W, H, stream, A.ws, ms.dp);
    if (st != RTG_OK) return st;
    ms.blocks = persistent_blocks(A);
    if (A.ws.grow(A.ws.spill, spill_ints(ms.blocks) * sizeof(int))) return fail(RTG_E_NOMEM, "spill");
    ms.tables = A.ws.tables.as<float>();
    ms.spill = A.ws.spill.as<int>();
    return enqueue_megakernel(S, p, W, H, o, stream, A, cancel, ms);
}

}  // namespace rtg
