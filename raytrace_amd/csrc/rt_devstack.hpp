// rt_devstack.hpp — traversal stack of the device kernels and the XCD-aware
// persistent work split shared by the megakernel and the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rt_scene.hpp"

namespace rtg {

constexpr int kBlockThreads = 128;  // two wave64s
constexpr int kLdsStack = 16;       // entries per thread in LDS (8 KB per block)
// A Karras tree over 62-bit unique keys is fewer than 64 binary levels deep,
// so its 4-wide form has at most 32 levels, each leaving at most 3 entries.
constexpr int kStackMax = 96;
static_assert(3 * kMaxLevels4 <= kStackMax, "the stack holds any walk of a tree the build emits");

// Per-thread stack: the first kLdsStack entries in LDS (thread-interleaved, no
// bank conflicts), the rest in a per-thread global spill area.
struct DevStack {
    int *lds;
    int *spill;
    int gstride;
    int sp;
    const Bvh4Node *top;  // LDS copy of the nodes below kTopNodes (one-scene kernels), or null
    __device__ void reset() { sp = 0; }
    // The LDS part is addressed as LDS (ds_read / ds_write); the compiler
    // barriers keep each LDS access from being merged with the spill access
    // of the other branch into one flat access, which would take the vector
    // memory path and its latency on every push and pop.
    __device__ void push(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
        typedef int __attribute__((address_space(3))) LdsInt;
        if (sp < kLdsStack) {
            ((LdsInt *)lds)[sp * kBlockThreads] = v;
            __asm__ volatile("" ::: "memory");
        } else {
            spill[(sp - kLdsStack) * gstride] = v;
        }
#endif
        ++sp;
    }
    __device__ int pop() {
        --sp;
        int v = 0;
#if defined(__HIP_DEVICE_COMPILE__)
        typedef const int __attribute__((address_space(3))) LdsInt;
        if (sp < kLdsStack) {
            v = ((LdsInt *)lds)[sp * kBlockThreads];
            __asm__ volatile("" ::: "memory");
        } else {
            v = spill[(sp - kLdsStack) * gstride];
        }
#endif
        return v;
    }
    __device__ bool empty() const { return sp == 0; }
    __device__ void on_node() {}
    __device__ void on_prim() {}
    // v holds on at least N/D of the wave's active lanes
    template <int N, int D>
    __device__ bool lanes_at_least(bool v) const {
        return D * __popcll(__ballot(v)) >= N * __popcll(__ballot(true));
    }
    __device__ Bvh4Node node(const SceneView &S, int i) const {
#if defined(__HIP_DEVICE_COMPILE__)
        typedef const Bvh4Node __attribute__((address_space(3))) LdsNode;  // `top` points into LDS
        if (top != nullptr && i < kTopNodes) {
            const Bvh4Node n = ((LdsNode *)top)[i];
            // a compiler barrier (no instruction): keeps this LDS read from
            // being merged with the HBM read below into one flat load
            __asm__ volatile("" ::: "memory");
            return n;
        }
#endif
        return S.nodes[i];
    }
};

// Copy the scene's top nodes into `dst` (a kTopNodes array in LDS) with the
// whole block, then let `st` read them from there.  Every thread of the
// block must call it (it synchronises the block).
__device__ inline void stack_top_nodes(DevStack &st, Bvh4Node *dst, const SceneView &S) {
    const int nt = min(kTopNodes, S.n_nodes);
    const float4 *src = reinterpret_cast<const float4 *>(S.nodes);
    float4 *d = reinterpret_cast<float4 *>(dst);
    for (int i = threadIdx.x; i < 4 * nt; i += blockDim.x) d[i] = src[i];
    __syncthreads();
    st.top = dst;
}

__device__ inline DevStack make_stack(int *lds_base, int *spill) {
    return DevStack{lds_base + threadIdx.x, spill + (size_t)blockIdx.x * kBlockThreads + threadIdx.x,
                    (int)gridDim.x * kBlockThreads, 0, nullptr};
}

// Spill entries needed for a persistent grid of `blocks` blocks.
inline size_t spill_ints(int blocks) { return (size_t)blocks * kBlockThreads * (kStackMax - kLdsStack); }

// XCD-aware split of `nchunks` work chunks over a persistent grid: block b
// runs on XCD b % 8 and takes chunks of that XCD's contiguous eighth, so
// neighbouring chunks share an L2.  Every wave of a block sees the same trip
// count.
struct ChunkRange {
    int first, end, step;
};
__device__ inline ChunkRange xcd_chunks(int nchunks) {
    const int G = gridDim.x, xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per = G >> 3;
    const int region = (nchunks + 7) >> 3;
    const int start = xcd * region;
    return ChunkRange{start + local, min(nchunks, start + region), per};
}

// Device counter alone in its 256-byte line: counters that many waves hit
// with atomics must not share a line (same-line atomics serialise).
struct alignas(256) LineCtr {
    unsigned v;
    unsigned pad[63];
};

// Per-wave dynamic work split: a wave grabs chunks of 64 << grab_lg items,
// first from its XCD's contiguous eighth of the range (one counter per XCD, so
// neighbouring chunks share an L2), then from the other XCDs' eighths once its
// own is exhausted, and hands a grabbed chunk out 64 items per next() call.
// `ctr` points to 8 counters that are zero at launch.  Every lane of the wave
// must call next(); it returns the first item of the next 64, or -1 when every
// chunk is taken.  A region whose counter was seen past its end is never tried
// again by the wave, so a wave makes at most 8 failed grabs.
struct WaveQueue {
    LineCtr *ctr;
    int n, glg, nchunks, region, x;
    int cur, end;   // the grabbed chunk: next item handed out, end (wave-uniform)
    uint32_t dead;  // regions this wave has seen exhausted
    __device__ WaveQueue(LineCtr *c, int n_items, int grab_lg = 0)
        : ctr(c), n(n_items), glg(grab_lg), nchunks((n_items + (64 << grab_lg) - 1) >> (6 + grab_lg)),
          region((nchunks + 7) >> 3), x(blockIdx.x & 7), cur(0), end(0), dead(0) {}
    __device__ int region_len(int r) const { return max(0, min(region, nchunks - r * region)); }
    __device__ int next() {
        if (cur + 64 < end) return cur += 64;
        if (nchunks == 0) return -1;
        const int lane = __lane_id();
        for (;;) {
            int c = 0;
            if (lane == 0) c = (int)atomicAdd(&ctr[x].v, 1u);
            c = __builtin_amdgcn_readfirstlane(c);  // wave-uniform: an SGPR
            if (c < region_len(x)) {
                cur = (x * region + c) << (6 + glg);
                end = min(n, cur + (64 << glg));
                return cur;
            }
            dead |= 1u << x;
            // regions with chunks left, read in one round trip
            bool left = false;
            if (lane < 8 && !((dead >> lane) & 1u))
                left = (int)__hip_atomic_load(&ctr[lane].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       region_len(lane);
            const uint32_t live = (uint32_t)__ballot(left) & 0xFFu & ~dead;
            if (live == 0) return -1;
            const uint32_t rot = (live >> x) | (live << (8 - x));  // bit i: region (x + i) & 7
            x = (x + __builtin_ctz(rot & 0xFFu)) & 7;
        }
    }
};

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Wave-aggregated append: every lane of the wave must call it (count 0 for
// inactive lanes).  Returns this lane's first slot in the queue.
__device__ inline int wave_append(unsigned *counter, int count) {
    const int lane = __lane_id();
    int x = count;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int total = __shfl(x, 63, 64);
    int base = 0;
    if (lane == 63 && total > 0) base = (int)atomicAdd(counter, (unsigned)total);
    base = __shfl(base, 63, 64);
    return base + x - count;
}

// Wave-aggregated append of 0 or 1 item per lane (ballot + bit count).
__device__ inline int wave_append1(unsigned *counter, bool take) {
    const unsigned long long m = __ballot(take);
    const int lane = __lane_id();
    int base = 0;
    if (lane == 0 && m != 0ull) base = (int)atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl(base, 0, 64);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

}  // namespace rtg
