// rt_devstack.hpp — traversal stack of the device kernels and the XCD-aware
// persistent work split shared by the megakernel and the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace rtg {

constexpr int kBlockThreads = 128;  // two wave64s
constexpr int kLdsStack = 32;       // entries per thread in LDS
constexpr int kStackMax = 64;       // a Karras tree over 62-bit unique keys is < 63 levels deep

// Per-thread stack: the first kLdsStack entries in LDS (thread-interleaved, no
// bank conflicts), the rest in a per-thread global spill area.
struct DevStack {
    int *lds;
    int *spill;
    int gstride;
    int sp;
    __device__ void reset() { sp = 0; }
    __device__ void push(int v) {
        if (sp < kLdsStack)
            lds[sp * kBlockThreads] = v;
        else
            spill[(sp - kLdsStack) * gstride] = v;
        ++sp;
    }
    __device__ int pop() {
        --sp;
        return sp < kLdsStack ? lds[sp * kBlockThreads] : spill[(sp - kLdsStack) * gstride];
    }
    __device__ bool empty() const { return sp == 0; }
};

__device__ inline DevStack make_stack(int *lds_base, int *spill) {
    return DevStack{lds_base + threadIdx.x, spill + (size_t)blockIdx.x * kBlockThreads + threadIdx.x,
                    (int)gridDim.x * kBlockThreads, 0};
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

}  // namespace rtg
