// rtg_internal.hpp — library-internal types of librtg (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
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
    DevBuf nodes, prims, objs, brute, parts, tri_nrm, tri_uv, tri_pos, tri_part, tri_mtl, tri_bins, mtls, lights, texs, texels;
    SceneView view{};
    rtg_camera camera{};
    uint64_t geom_hash = 0;
    double build_ms = 0.0;
    size_t bytes = 0;
    int n_nodes = 0;
    uint32_t lights_on = 0;  // bit i: light i is switched on
    int max_fanout = 2;      // most child rays per hit over the shown objects' materials
    uint64_t rebuilds = 0;         // rtg_scene_update calls that rebuilt the BVH
    uint64_t refits = 0;           // rtg_scene_update calls that refitted the BVH
    uint64_t shading_updates = 0;  // rtg_scene_update calls that only re-uploaded shading tables
    // Tree topology of the last full build, kept for refits: sorted keys
    // (primitive id in the low 32 bits), radix-tree children, leaf ranges,
    // parents and emitted-node indices of the Karras tree over topo_n
    // primitives; topo_hash identifies the primitive set and order.
    DevBuf t_keys, t_left, t_right, t_first, t_last, t_pint, t_pleaf, t_newidx;
    int topo_n = 0, topo_emit = 0;
    uint64_t topo_hash = 0;
};

// Device memory of one in-flight render (or batch): traversal spill stacks,
// ray-generation tables, wavefront queues and counters, batch tables.  A
// workspace serves one render at a time; concurrent renders (serving slots)
// each have their own.
struct Workspace {
    static constexpr int kLevels = RTG_MAX_LEVEL + 1;
    DevBuf spill, tables;
    uint64_t tables_key = 0;
    DevBuf w_rays[kLevels], w_hits[kLevels], w_cs[kLevels], w_nodes[kLevels], w_masks[kLevels],
        w_slots[kLevels], w_shadow[kLevels], w_samples, w_qcount;
    DevBuf b_reqs, b_scenes, b_chunks;  // batch: request table, scene table, chunk -> request
    unsigned *qcount_host = nullptr;    // pinned
    void *staging = nullptr;            // pinned upload staging (batch tables)
    size_t staging_bytes = 0;
    uint64_t epoch = 0;  // bumped when a buffer moves (captured graphs hold addresses)
    hipGraphExec_t graph = nullptr;
    uint64_t graph_key = 0;
    static constexpr int kSeen = 32;
    uint64_t seen_keys[kSeen] = {};  // batch shapes launched recently, with launch counts
    uint32_t seen_count[kSeen] = {};
    int seen_next = 0;
    // Recorded after the last render enqueued with this workspace: a render
    // on another stream waits for it (device-side) before reusing the buffers.
    hipEvent_t done = nullptr;
    int order_before(hipStream_t s);  // make `s` wait for the previous render
    int mark_done(hipStream_t s);     // record `done` on `s`
    // Side stream of the level loop: the shadow rays of level k run on it
    // while the render stream traces level k + 1, and the last level's
    // shadow rays run on the render stream beside it (own spill area: two
    // traversal kernels at once).  Created on first use, default priority.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork[kLevels] = {}, ev_join = nullptr, ev_join2 = nullptr;
    DevBuf spill2, spill3;  // spill areas of the side stream and of the second side stream
    int side_init();
    // Grow `b` to hold `bytes`; bumps the epoch if it moved.
    int grow(DevBuf &b, size_t bytes);
    void release();  // frees everything (the stream using it must be idle)
    ~Workspace() { release(); }
};

// One concurrent batch of the serving loop: its stream, workspace, device
// outputs and pinned host staging, and the requests it carries.
struct ServeSlot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    DevBuf out, counts;
    uint8_t *pinned = nullptr;
    size_t pinned_bytes = 0;
    unsigned long long *counts_host = nullptr;  // pinned, 4
    std::vector<int> reqs;
    std::vector<size_t> offs;
    bool busy = false;
    ServeSlot() = default;
    ServeSlot(const ServeSlot &) = delete;
    ServeSlot &operator=(const ServeSlot &) = delete;
    ~ServeSlot();
};

// Per-device render state: default stream and workspace, outputs, cancel
// flag, kernel timing, serving slots.
// Launch tuning of the wavefront engine (rtg_tuning): pixel block of a wave
// (tile_lg = log2 of its width, 2..5; height 64 >> tile_lg), 64-item queue
// chunks per grab (1 << grab_lg, 0..2), persistent grid in percent of the
// resident blocks (25..100).
struct Tuning {
    int tile_lg = 3, grab_lg = 0, grid_pct = 100;
    uint64_t key() const { return (uint64_t)tile_lg | ((uint64_t)grab_lg << 4) | ((uint64_t)grid_pct << 8); }
};

struct Arena {
    static constexpr int kLevels = Workspace::kLevels;
    Tuning tune;
    int device = -1;
    int cus = 0;
    // Largest queue memory (bytes) one wavefront may reserve to run without
    // host round trips between levels: a quarter of HBM, 2..64 GiB.
    size_t no_sync_budget = size_t(8) << 30;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf counters, out_rgb, out_ids, out_t;
    // Pinned staging of a batch's results: one device-to-host copy per batch,
    // then a threaded scatter into the callers' arrays.
    uint8_t *pinned_out = nullptr;
    size_t pinned_out_bytes = 0;
    Workspace ws;
    std::vector<std::unique_ptr<ServeSlot>> serve_slots;
    int *cancel_host = nullptr;  // pinned, mapped: the cancel flag of the job rendering now
    int *cancel_dev = nullptr;
    std::mutex cancel_lock;        // guards cancel_job and writes of *cancel_host
    const void *cancel_job = nullptr;  // the asynchronous job the flag belongs to, or null
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

enum KernelKind { KT_EXTEND, KT_SHADE, KT_OCCLUDE, KT_COMBINE, KT_RESOLVE, KT_MEGA, KT_LOCAL, KT_COUNT };
// Start timing a launch of `kind` on `s` (no-op unless A.ktimes); returns a
// record index for ktime_end, or -1.
int ktime_begin(Arena &A, hipStream_t s, int kind, uint64_t items);
// End the record; `dev_items` (optional) is a device counter whose value at
// this point of the stream is the launch's item count.
void ktime_end(Arena &A, hipStream_t s, int rec, const unsigned *dev_items = nullptr);

int upload_raygen_tables(const SceneRes &S, int W, int H, hipStream_t stream, Workspace &ws, double &dp);
// Ray-generation tables of a W x H frame (RayTracer.cpp:13-27): a[x] =
// (float)((x - W/2) * dp) then b[y] likewise, dp = tan(fovy*PI/360)/(H/2).
double raygen_tables(float fovy, int W, int H, float *tab);
int persistent_blocks(const Arena &A);

Arena *arena();
int ensure(DevBuf &b, size_t bytes);

// Build S from d.  With `donor` holding the topology of a tree over the same
// primitive set (same count and order), that topology moves into S and the
// tree is refitted: primitives and boxes are recomputed, the sort and the
// radix tree are skipped.
int scene_build(SceneRes &S, const rtg_scene_desc &d, hipStream_t stream, SceneRes *donor = nullptr);
// Update of a scene whose geometry is unchanged: shading tables only.
int scene_update_shading(SceneRes &S, const rtg_scene_desc &d, hipStream_t stream);
// Frame size of a render (params or the scene camera's) and parameter checks.
int resolve_size(const SceneRes &S, const rtg_params *p, int &W, int &H);


struct RenderOut {
    uint8_t *rgb;   // device
    int32_t *ids;   // device, optional
    float *tprim;   // device, optional
    unsigned long long *counts;  // device, 4 counters (accumulated)
    unsigned long long *retests = nullptr;  // device, 1 counter: FP64 re-tests (optional)
};

// Cooperative cancel of a frame (rtg_cancel of an asynchronous job).  `host`
// and `dev` are the host and device addresses of one mapped flag.  Kernels
// poll `dev`; the host checks `host` between levels and passes.  A wavefront
// frame that sees it stops in its current pass and leaves that pass's pixels
// untouched; passes of at most `pass_samples` level-0 samples (0: no limit
// beyond the engine's own) make the cancel take effect part-way through a
// frame, as the reference's workers stop between pixels.
struct Cancel {
    const int *host;
    const int *dev;
    size_t pass_samples;
};

// Launch one frame of the megakernel on `stream` (asynchronous); `cancel`
// may be null.
int launch_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                      Arena &A, const Cancel *cancel);
// The megakernel's scratch: ray-generation tables (W + H floats, written for
// the frame's W, H and `dp`), the traversal stacks' spill area for `blocks`
// persistent blocks (spill_ints(blocks) ints).
struct MegaScratch {
    float *tables;
    int *spill;
    int blocks;
    double dp;
};
// Enqueue the megakernel with the given scratch (tables already written);
// allocates nothing.
int enqueue_megakernel(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                       Arena &A, const Cancel *cancel, const MegaScratch &ms);
// Enqueue the ray-generation tables (float)((x - W/2) dp), (float)((y - H/2) dp)
// into W + H floats at `tab` on `stream`; allocates nothing.
int write_raygen_tables(float *tab, int W, int H, double dp, hipStream_t stream);
// Set up the no-allocation entry point's scratch on `device` (current).
int noalloc_setup(int device);
// Launch one frame of the wavefront engine on `stream` (asynchronous) with
// the buffers of `ws`; `cancel` may be null.
// Host copies of device results staged in pinned memory, run on up to 8
// host threads in pieces of at most 1 MB.
struct HostCopy {
    void *dst;
    const uint8_t *src;
    size_t bytes;
};
void copy_out(const HostCopy *cp, int n);

// Grow `ws` to what one W x H frame of the given depth, fanout and lights
// reserves before it launches (nothing is allocated while it renders).
int reserve_frame(Workspace &ws, int W, int H, int spp, int max_level, int fanout, int n_on, Arena &A);
// Grow `ws` to what a batch of up to n_reqs requests over n_scenes scenes,
// `samples` level-0 samples and `table_floats` ray-generation table entries
// reserves before it launches.
int reserve_batch(Workspace &ws, size_t samples, int n_reqs, int n_scenes, size_t table_floats, int max_level,
                  int fanout, int n_on, hipStream_t stream, Arena &A);
int launch_wavefront(const SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                     Workspace &ws, Arena &A, const Cancel *cancel);

// One request of a batch (rtg_render_batch): a frame of `scene` into device
// outputs.
struct BatchItem {
    const SceneRes *scene;
    rtg_params params;
    int W, H;
    uint8_t *rgb;   // device
    int32_t *ids;   // device, optional
    float *tprim;   // device, optional
};
// Enqueue every request of a batch on `stream` as one wavefront: all their
// rays share one queue per level; no host round trip when the queue bounds fit
// the workspace budget.  With capture_at = k > 0 the launch sequence of a
// padded batch shape is captured into a hipGraph at its k-th launch and
// replayed afterwards (shapes seen fewer times run eagerly); 0: never.
// `counts` (device, 4 x u64) accumulates the ray counts of the whole batch.
int launch_batch(const BatchItem *items, int n, unsigned long long *counts, unsigned long long *retests,
                 hipStream_t stream, Workspace &ws, Arena &A, int capture_at, int *graph_replayed);
// Device bytes the queues of a batch reserve (its no-sync bound), for
// admission by memory budget; 0 when the bound exceeds `budget`.
size_t batch_queue_bytes(size_t samples, int max_level, int fanout, int n_on, size_t budget);
// Level-0 queue length of a single frame.
// Level-0 samples of a W x H frame rendered in pixel blocks of (1 << tile_lg) x (64 >> tile_lg).
inline size_t frame_samples(int W, int H, int spp, int tile_lg) {
    const int tw = 1 << tile_lg, th = 64 >> tile_lg;
    return (size_t)((W + tw - 1) / tw) * ((H + th - 1) / th) * 64 * (spp > 0 ? spp : 1);
}

}  // namespace rtg

// Opaque scene handle of the C ABI.
struct rtg_scene_s {
    rtg::SceneRes res;
};
