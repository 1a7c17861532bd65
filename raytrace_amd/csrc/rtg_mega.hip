// rtg_mega.hip — the one-launch-per-frame megakernel.
//
// Persistent grid (a few blocks per CU), 128 threads per block = two waves,
// each wave an 8x8 pixel block.  Pixel blocks are split into 8 contiguous
// image regions, one per XCD (block b runs on XCD b % 8), so each XCD's L2
// caches the part of the BVH its region touches.  Traversal stacks live in
// LDS (32 entries per thread) and spill to a per-thread global area; a Karras
// tree over 62-bit unique keys is at most 63 levels deep, so 64 entries bound
// the stack.
#include "rtg_internal.hpp"

namespace rtg {

constexpr int kMegaBlock = 128;
constexpr int kLdsStack = 32;
constexpr int kStackMax = 64;

struct DevStack {
    int *lds;
    int *spill;
    int gstride;
    int sp;
    __device__ void push(int v) {
        if (sp < kLdsStack)
            lds[sp * kMegaBlock] = v;
        else
            spill[(sp - kLdsStack) * gstride] = v;
        ++sp;
    }
    __device__ int pop() {
        --sp;
        return sp < kLdsStack ? lds[sp * kMegaBlock] : spill[(sp - kLdsStack) * gstride];
    }
    __device__ bool empty() const { return sp == 0; }
};

struct MegaArgs {
    int W, H, spp, max_level, tile_compat;
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

__device__ inline uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kMegaBlock) void k_megakernel(SceneView S, MegaArgs A) {
    __shared__ int lds[kLdsStack * kMegaBlock];
    const int tid = threadIdx.x;
    DevStack st{lds + tid, A.spill + (size_t)blockIdx.x * kMegaBlock + tid, (int)gridDim.x * kMegaBlock, 0};
    RayCounts rc{0, 0, 0, 0};
    const int G = gridDim.x, xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per_xcd = G >> 3;
    const int region = (A.nblocks + 7) >> 3;
    const int start = xcd * region, end = min(A.nblocks, start + region);
    const int tw = (A.W >> 6) << 6, th = (A.H >> 6) << 6;
    for (int pb = start + local; pb < end; pb += per_xcd) {
        if (__builtin_nontemporal_load(A.cancel) != 0) break;
        const int bx = pb % A.bw, by = pb / A.bw;
        const int px = bx * 16 + ((tid >> 6) << 3) + (tid & 7), py = by * 8 + ((tid & 63) >> 3);
        if (px >= A.W || py >= A.H) continue;
        if (A.tile_compat && (px >= tw || py >= th)) continue;
        V3 acc = V3{0, 0, 0};
        PrimaryOut po{-1, -1, kMiss};
        for (int s = 0; s < A.spp; ++s) {
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
            PrimaryOut p0;
            const V3 c = render_sample(S, A.max_level, q, st, rc, p0);
            if (s == 0) po = p0;
            acc = acc + c;
        }
        if (A.spp > 1) acc = acc * (1.0f / (float)A.spp);
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
    const uint32_t c0 = wave_sum(rc.primary), c1 = wave_sum(rc.shadow), c2 = wave_sum(rc.reflect),
                   c3 = wave_sum(rc.refract);
    if ((tid & 63) == 0) {
        atomicAdd(&A.counts[0], (unsigned long long)c0);
        atomicAdd(&A.counts[1], (unsigned long long)c1);
        atomicAdd(&A.counts[2], (unsigned long long)c2);
        atomicAdd(&A.counts[3], (unsigned long long)c3);
    }
}

int launch_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                      Arena &A) {
    // ray-generation tables: (float)(xcur * dp), (float)(ycur * dp) (RayTracer.cpp:14,26)
    const double dp = tan((double)S.camera.fovy * 3.1415926535897932384 / 360) / (H / 2);
    std::vector<float> tab((size_t)W + H);
    for (int x = 0; x < W; ++x) tab[x] = (float)((x - W / 2) * dp);
    for (int y = 0; y < H; ++y) tab[W + y] = (float)((y - H / 2) * dp);
    if (ensure(A.tables, tab.size() * sizeof(float))) return fail(RTG_E_NOMEM, "tables");
    RTG_HIP(hipMemcpyAsync(A.tables.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    const int blocks = std::max(8, (A.cus > 0 ? A.cus : 256) * 8);
    const size_t spill = (size_t)blocks * kMegaBlock * (kStackMax - kLdsStack) * sizeof(int);
    if (ensure(A.spill, spill)) return fail(RTG_E_NOMEM, "spill");
    MegaArgs a;
    a.W = W, a.H = H, a.spp = p.spp > 0 ? p.spp : 1, a.max_level = p.max_level, a.tile_compat = p.tile_compat;
    a.bw = (W + 15) / 16;
    a.nblocks = a.bw * ((H + 7) / 8);
    a.dpf = (float)dp;
    a.seed = p.seed;
    a.ax = A.tables.as<float>();
    a.by = A.tables.as<float>() + W;
    a.rgb = o.rgb, a.ids = o.ids, a.tprim = o.tprim;
    a.counts = o.counts;
    a.cancel = A.cancel_dev;
    a.spill = A.spill.as<int>();
    k_megakernel<<<blocks, kMegaBlock, 0, stream>>>(S.view, a);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

}  // namespace rtg
