// rtg_internal.hpp — library-internal types of librtg (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rtg.h"
#include "rt_shade.hpp"

namespace rtg {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
#define RTG_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(RTG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device allocation owned by a scene or the render arena.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int alloc(size_t n);  // RTG_OK / RTG_E_NOMEM
    void release();
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

struct SceneRes {
    int device = 0;
    DevBuf nodes, prims, objs, brute, parts, tri_nrm, tri_uv, tri_part, tri_bins, mtls, lights, texs, texels;
    SceneView view{};
    rtg_camera camera{};
    uint64_t geom_hash = 0;
    double build_ms = 0.0;
    size_t bytes = 0;
    int n_nodes = 0;
    uint32_t lights_on = 0;  // bit i: light i is switched on
    int max_fanout = 2;      // most child rays per hit over the shown objects' materials
};

// Per-device render state: stream, spill stacks, tables, counters, outputs.
struct Arena {
    int device = -1;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf spill, tables, counters, out_rgb, out_ids, out_t;
    uint64_t tables_key = 0;
    int *cancel_host = nullptr;  // pinned, mapped
    int *cancel_dev = nullptr;
    // wavefront engine: per-level queues and records, grown on demand
    static constexpr int kLevels = RTG_MAX_LEVEL + 1;
    DevBuf w_rays[kLevels], w_hits[kLevels], w_nodes[kLevels], w_masks[kLevels], w_slots[kLevels],
        w_shadow[kLevels], w_samples, w_qcount;
    unsigned *qcount_host = nullptr;  // pinned
    // kernel timing (RTG_FLAG_KERNEL_TIMES)
    struct KernelRec {
        int kind;
        hipEvent_t e0, e1;
        uint64_t items;
        int slot;  // >= 0: item count read back from the device into kitems_host[slot]
    };
    bool ktimes = false;
    std::vector<KernelRec> krecs;
    std::vector<hipEvent_t> kevents;  // free timing events
    unsigned *kitems_host = nullptr;  // pinned
    int kitems_next = 0;
    static constexpr int kItemSlots = 4096;
    std::mutex lock;  // one render at a time on the arena
};

enum KernelKind { KT_EXTEND, KT_SHADE, KT_OCCLUDE, KT_COMBINE, KT_RESOLVE, KT_MEGA, KT_COUNT };
// Start timing a launch of `kind` on `s` (no-op unless A.ktimes); returns a
// record index for ktime_end, or -1.
int ktime_begin(Arena &A, hipStream_t s, int kind, uint64_t items);
// End the record; `dev_items` (optional) is a device counter whose value at
// this point of the stream is the launch's item count.
void ktime_end(Arena &A, hipStream_t s, int rec, const unsigned *dev_items = nullptr);

int upload_raygen_tables(const SceneRes &S, int W, int H, hipStream_t stream, Arena &A, double &dp);
int persistent_blocks(const Arena &A);

Arena *arena();
int ensure(DevBuf &b, size_t bytes);

int scene_build(SceneRes &S, const rtg_scene_desc &d, hipStream_t stream);

struct RenderOut {
    uint8_t *rgb;   // device
    int32_t *ids;   // device, optional
    float *tprim;   // device, optional
    unsigned long long *counts;  // device, 4 counters (accumulated)
};

// Launch one frame of the megakernel on `stream` (asynchronous).
int launch_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                      Arena &A);
// Launch one frame of the wavefront engine on `stream` (asynchronous).
int launch_wavefront(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                     Arena &A);

}  // namespace rtg
