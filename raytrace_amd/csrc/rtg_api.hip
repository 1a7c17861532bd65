// rtg_api.hip — implementation of the librtg C ABI (include/rtg.h).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>

#include "rt_prep.hpp"
#include "rtg_internal.hpp"

namespace rtg {

namespace {
thread_local std::string g_err = "";
std::mutex g_init_lock;
// The device state of rtg_init.  Owned by rtg_init / rtg_shutdown only: it is
// never destroyed by a static destructor at process exit, where the HIP
// runtime (or a profiler attached to it) may already be torn down and a
// stream or event call aborts; the driver reclaims the memory with the
// process.
Arena *g_arena = nullptr;


uint64_t hash_bytes(uint64_t h, const void *p, size_t n) {
    const uint8_t *c = static_cast<const uint8_t *>(p);
    size_t i = 0;
    uint64_t lanes[4] = {h ^ 0x9E3779B97F4A7C15ull, h + 0xC2B2AE3D27D4EB4Full, h * 3 + 1, ~h};
    for (; i + 32 <= n; i += 32)
        for (int k = 0; k < 4; ++k) {
            uint64_t w;
            memcpy(&w, c + i + 8 * k, 8);
            lanes[k] = (lanes[k] ^ w) * 0x100000001B3ull;
            lanes[k] ^= lanes[k] >> 29;
        }
    uint64_t t = n;
    for (; i < n; ++i) t = t * 131 + c[i];
    return mix64(lanes[0] ^ mix64(lanes[1] ^ mix64(lanes[2] ^ mix64(lanes[3] ^ t))));
}

// Hash of everything the BVH depends on.
uint64_t geometry_hash(const rtg_scene_desc &d) {
    uint64_t h = hash_bytes(0, d.objects, sizeof(rtg_object) * (size_t)d.n_objects);
    for (int i = 0; i < d.n_parts; ++i) {
        const rtg_part &p = d.parts[i];
        h = hash_bytes(h, &p.tri_count, sizeof(int32_t) * 2);
        h = hash_bytes(h, &p.bmin, sizeof(rtg_vec4) * 2);
        if (p.tris && p.tri_count > 0) h = hash_bytes(h, p.tris, sizeof(rtg_triangle) * (size_t)p.tri_count);
    }
    return h;
}

// Primitive set and order of a scene (what a BVH topology depends on): each
// object's type, visibility and part range, and each part's triangle count.
uint64_t topology_hash(const rtg_scene_desc &d) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)d.n_objects;
    for (int i = 0; i < d.n_objects; ++i) {
        const rtg_object &o = d.objects[i];
        const int32_t v[4] = {o.type, o.show, o.part_begin, o.part_count};
        h = hash_bytes(h, v, sizeof(v));
    }
    for (int i = 0; i < d.n_parts; ++i) h = hash_bytes(h, &d.parts[i].tri_count, sizeof(int32_t));
    return h;
}

int validate_desc(const rtg_scene_desc *d) {
    if (d == nullptr) return fail(RTG_E_INVALID, "scene descriptor is NULL");
    if (d->n_objects < 0 || d->n_parts < 0 || d->n_materials < 0 || d->n_lights < 0 || d->n_textures < 0)
        return fail(RTG_E_INVALID, "negative count in scene descriptor");
    if ((d->n_objects && !d->objects) || (d->n_parts && !d->parts) || (d->n_materials && !d->materials) ||
        (d->n_lights && !d->lights) || (d->n_textures && !d->textures))
        return fail(RTG_E_INVALID, "NULL array in scene descriptor");
    if (d->n_lights > RTG_MAX_LIGHTS) return fail(RTG_E_LIMIT, "more than RTG_MAX_LIGHTS lights");
    if (d->n_materials > (1 << 20)) return fail(RTG_E_LIMIT, "more than 2^20 materials");
    for (int i = 0; i < d->n_materials; ++i)
        if (d->materials[i].tex >= d->n_textures) return fail(RTG_E_INVALID, "material texture out of range");
    for (int i = 0; i < d->n_parts; ++i)
        if (d->parts[i].tri_count > 32767) return fail(RTG_E_LIMIT, "part with more than 32767 triangles");
    return RTG_OK;
}


}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int resolve_size(const SceneRes &S, const rtg_params *p, int &W, int &H) {
    if (p == nullptr) return fail(RTG_E_INVALID, "params is NULL");
    W = p->width > 0 ? p->width : S.camera.width;
    H = p->height > 0 ? p->height : S.camera.height;
    if (W <= 0 || H <= 0 || W > RTG_MAX_DIM || H > RTG_MAX_DIM) return fail(RTG_E_LIMIT, "frame size out of range");
    if (p->max_level < 0 || p->max_level > RTG_MAX_LEVEL) return fail(RTG_E_LIMIT, "max_level out of range");
    if (!(p->spp == 0 || p->spp == 1 || p->spp == 4)) return fail(RTG_E_INVALID, "spp must be 1 or 4");
    return RTG_OK;
}

__global__ void k_fill_u8(uint8_t *p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

int DevBuf::alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    if (hipMalloc(&p, n) != hipSuccess) {
        p = nullptr;
        return RTG_E_NOMEM;
    }
    bytes = n;
    return RTG_OK;
}
void DevBuf::release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
int ensure(DevBuf &b, size_t n) {
    if (b.p != nullptr && b.bytes >= n) return RTG_OK;
    return b.alloc(n);
}

int Workspace::grow(DevBuf &b, size_t n) {
    if (b.p != nullptr && b.bytes >= n) return RTG_OK;
    ++epoch;
    return b.alloc(n + n / 4);  // headroom: fewer moves, fewer graph re-captures
}

int Workspace::order_before(hipStream_t s) {
    if (done != nullptr) RTG_HIP(hipStreamWaitEvent(s, done, 0));
    return RTG_OK;
}

int Workspace::mark_done(hipStream_t s) {
    if (done == nullptr) RTG_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    RTG_HIP(hipEventRecord(done, s));
    return RTG_OK;
}

int Workspace::side_init() {
    if (side == nullptr) RTG_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    for (int k = 0; k < kLevels; ++k)
        if (ev_fork[k] == nullptr) RTG_HIP(hipEventCreateWithFlags(&ev_fork[k], hipEventDisableTiming));
    if (ev_join == nullptr) RTG_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    if (ev_join2 == nullptr) RTG_HIP(hipEventCreateWithFlags(&ev_join2, hipEventDisableTiming));
    return RTG_OK;
}

void Workspace::release() {
    if (side) (void)hipStreamSynchronize(side);
    for (int k = 0; k < kLevels; ++k) {
        if (ev_fork[k]) (void)hipEventDestroy(ev_fork[k]);
        ev_fork[k] = nullptr;
    }
    if (ev_join) (void)hipEventDestroy(ev_join);
    ev_join = nullptr;
    if (ev_join2) (void)hipEventDestroy(ev_join2);
    ev_join2 = nullptr;
    if (side) (void)hipStreamDestroy(side);
    side = nullptr;
    spill2.release();
    spill3.release();
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    if (done) (void)hipEventDestroy(done);
    done = nullptr;
    spill.release(), tables.release(), w_samples.release(), w_qcount.release();
    b_reqs.release(), b_scenes.release(), b_chunks.release();
    for (int k = 0; k < kLevels; ++k) {
        w_rays[k].release(), w_hits[k].release(), w_nodes[k].release();
        w_masks[k].release(), w_slots[k].release(), w_shadow[k].release(), w_cs[k].release();
    }
    if (qcount_host) (void)hipHostFree(qcount_host);
    if (staging) (void)hipHostFree(staging);
    qcount_host = nullptr;
    staging = nullptr;
    staging_bytes = 0;
    ++epoch;
}

Arena *arena() { return g_arena; }

static hipEvent_t take_event(Arena &A) {
    if (!A.kevents.empty()) {
        hipEvent_t e = A.kevents.back();
        A.kevents.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int ktime_begin(Arena &A, hipStream_t s, int kind, uint64_t items) {
    if (!A.ktimes) return -1;
    Arena::KernelRec r{kind, take_event(A), take_event(A), items, -1};
    if (r.e0 == nullptr || r.e1 == nullptr || hipEventRecord(r.e0, s) != hipSuccess) {
        if (r.e0) A.kevents.push_back(r.e0);
        if (r.e1) A.kevents.push_back(r.e1);
        return -1;
    }
    A.krecs.push_back(r);
    return (int)A.krecs.size() - 1;
}

void ktime_end(Arena &A, hipStream_t s, int rec, const unsigned *dev_items) {
    if (rec < 0) return;
    Arena::KernelRec &r = A.krecs[rec];
    (void)hipEventRecord(r.e1, s);
    if (dev_items != nullptr && A.kitems_host != nullptr && A.kitems_next < Arena::kItemSlots) {
        r.slot = A.kitems_next++;
        (void)hipMemcpyAsync(A.kitems_host + r.slot, dev_items, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    }
}

// Render one frame into device outputs; the caller holds A.lock.  `cancel`:
// the job's cancel flag (asynchronous jobs), else null.
static int render_device(SceneRes &S, const rtg_params &p, int W, int H, const RenderOut &o, hipStream_t stream,
                         Arena &A, const Cancel *cancel) {
    const size_t npix = (size_t)W * H;
    A.ktimes = (p.flags & RTG_FLAG_KERNEL_TIMES) != 0;
    if (A.ktimes && A.kitems_host == nullptr)
        RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&A.kitems_host), Arena::kItemSlots * sizeof(unsigned),
                              hipHostMallocDefault));
    k_fill_u8<<<1024, 256, 0, stream>>>(o.rgb, npix * 3, 127);
    if (o.ids) k_fill_u8<<<1024, 256, 0, stream>>>(reinterpret_cast<uint8_t *>(o.ids), npix * 8, 0xff);
    k_zero_u64<<<1, 64, 0, stream>>>(o.counts, 4);
    if (o.retests) k_zero_u64<<<1, 64, 0, stream>>>(o.retests, 1);
    if (!(p.mode == 0 || (p.mode >= RTG_MODE_CHECK && p.mode <= RTG_MODE_REFRACTTEST) || p.mode == RTG_MODE_RAYTRACE))
        return fail(RTG_E_INVALID, "unknown render mode");
    // AUTO: the wavefront engine.  The megakernel is slower on every large or
    // deep frame (its per-thread ray trees diverge); it only wins on small,
    // shallow frames that take well under a millisecond either way.  The
    // primary-only debug modes (checker, depth, normal, texture, material,
    // shadow) run in the megakernel: one closest hit per sample, no tree.
    int engine = p.engine == RTG_ENGINE_AUTO ? RTG_ENGINE_WAVEFRONT : p.engine;
    if (p.mode >= RTG_MODE_CHECK && p.mode <= RTG_MODE_SHADOWTEST) {
        if (p.flags & RTG_FLAG_FP64_RETEST) return fail(RTG_E_INVALID, "the FP64 re-test applies to ray-traced modes");
        engine = RTG_ENGINE_MEGAKERNEL;
    }
    if (engine == RTG_ENGINE_MEGAKERNEL && (p.flags & RTG_FLAG_FP64_RETEST))
        return fail(RTG_E_INVALID, "the FP64 near-tie re-test needs the wavefront engine");
    int st = A.ws.order_before(stream);
    if (st != RTG_OK) return st;
    st = engine == RTG_ENGINE_WAVEFRONT ? launch_wavefront(S, p, W, H, o, stream, A.ws, A, cancel)
                                        : launch_megakernel(S, p, W, H, o, stream, A, cancel);
    const int st2 = A.ws.mark_done(stream);
    return st != RTG_OK ? st : st2;
}

}  // namespace rtg

using namespace rtg;

struct rtg_job_s {
    std::thread th;
    std::atomic<int> status{RTG_E_PENDING};
    std::atomic<int> cancel{0};
    rtg_stats stats{};
};

extern "C" {

const char *rtg_last_error(void) { return g_err.c_str(); }
int rtg_abi_version(void) { return RTG_ABI_VERSION; }

int rtg_init(int device) {
    std::lock_guard<std::mutex> g(g_init_lock);
    if (g_arena && g_arena->device == device) return RTG_OK;
    if (g_arena) return fail(RTG_E_STATE, "librtg already initialised on another device");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(RTG_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(RTG_E_INVALID, "device index out of range");
    RTG_HIP(hipSetDevice(device));
    int st = noalloc_setup(device);
    if (st != RTG_OK) return st;
    std::unique_ptr<Arena> A(new Arena());
    A->device = device;
    hipDeviceProp_t prop;
    RTG_HIP(hipGetDeviceProperties(&prop, device));
    A->cus = prop.multiProcessorCount;
    A->no_sync_budget = std::min<size_t>(size_t(64) << 30, std::max<size_t>(size_t(2) << 30, prop.totalGlobalMem / 4));
    RTG_HIP(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
    RTG_HIP(hipEventCreate(&A->ev0));
    RTG_HIP(hipEventCreate(&A->ev1));
    RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&A->cancel_host), sizeof(int), hipHostMallocMapped));
    *A->cancel_host = 0;
    RTG_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&A->cancel_dev), A->cancel_host, 0));
    // ray counts [0..3] (primary, shadow, reflection, refraction), FP64 re-tests [4]
    if (ensure(A->counters, 8 * sizeof(unsigned long long))) return fail(RTG_E_NOMEM, "counters");
    g_arena = A.release();
    return RTG_OK;
}

int rtg_shutdown(void) {
    std::lock_guard<std::mutex> g(g_init_lock);
    if (!g_arena) return RTG_OK;
    Arena &A = *g_arena;
    {
        std::lock_guard<std::mutex> l(A.lock);
        (void)hipStreamSynchronize(A.stream);
        A.serve_slots.clear();
        A.ws.release();
        A.counters.release();
        A.out_rgb.release();
        A.out_ids.release();
        A.out_t.release();
        (void)hipHostFree(A.cancel_host);
        if (A.kitems_host) (void)hipHostFree(A.kitems_host);
        if (A.pinned_out) (void)hipHostFree(A.pinned_out);
        A.pinned_out = nullptr;
        A.pinned_out_bytes = 0;
        for (const Arena::KernelRec &r : A.krecs) A.kevents.push_back(r.e0), A.kevents.push_back(r.e1);
        for (hipEvent_t e : A.kevents) (void)hipEventDestroy(e);
        A.krecs.clear();
        A.kevents.clear();
        (void)hipEventDestroy(A.ev0);
        (void)hipEventDestroy(A.ev1);
        (void)hipStreamDestroy(A.stream);
    }
    delete g_arena;
    g_arena = nullptr;
    return RTG_OK;
}

int rtg_scene_create(const rtg_scene_desc *desc, rtg_scene_t *out) {
    if (out == nullptr) return fail(RTG_E_INVALID, "out is NULL");
    *out = nullptr;
    int st = validate_desc(desc);
    if (st != RTG_OK) return st;
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::unique_ptr<rtg_scene_s> s(new rtg_scene_s());
    s->res.device = A.device;
    {
        std::lock_guard<std::mutex> l(A.lock);
        st = scene_build(s->res, *desc, A.stream);
    }
    if (st != RTG_OK) return st;
    s->res.geom_hash = geometry_hash(*desc);
    s->res.topo_hash = topology_hash(*desc);
    *out = s.release();
    return RTG_OK;
}

// Update a resident scene: shading tables only when the geometry is the same;
// a BVH refit when only positions or shapes moved (same primitive set and
// order) and `allow_refit`; a full rebuild otherwise.
static int scene_update(rtg_scene_t scene, const rtg_scene_desc *desc, bool allow_refit) {
    if (scene == nullptr) return fail(RTG_E_INVALID, "scene is NULL");
    int st = validate_desc(desc);
    if (st != RTG_OK) return st;
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    // Same geometry (objects, parts, triangles): only materials, lights,
    // textures and the camera can have changed, so the BVH stays.
    const uint64_t gh = geometry_hash(*desc), th = topology_hash(*desc);
    SceneRes &R = scene->res;
    if (allow_refit && R.geom_hash != 0 && gh == R.geom_hash) {
        ++R.shading_updates;
        return scene_update_shading(R, *desc, A.stream);
    }
    const bool refit = allow_refit && R.topo_n > 0 && th == R.topo_hash;
    SceneRes fresh;
    fresh.device = A.device;
    st = scene_build(fresh, *desc, A.stream, refit ? &R : nullptr);
    if (st != RTG_OK) return st;
    auto take = [](DevBuf &a, DevBuf &b) { std::swap(a.p, b.p), std::swap(a.bytes, b.bytes); };
    take(R.nodes, fresh.nodes), take(R.prims, fresh.prims), take(R.objs, fresh.objs), take(R.brute, fresh.brute);
    take(R.parts, fresh.parts), take(R.tri_nrm, fresh.tri_nrm), take(R.tri_uv, fresh.tri_uv);
    take(R.tri_pos, fresh.tri_pos), take(R.tri_part, fresh.tri_part), take(R.tri_bins, fresh.tri_bins);
    take(R.tri_mtl, fresh.tri_mtl);
    take(R.mtls, fresh.mtls), take(R.lights, fresh.lights), take(R.texs, fresh.texs), take(R.texels, fresh.texels);
    take(R.t_keys, fresh.t_keys), take(R.t_left, fresh.t_left), take(R.t_right, fresh.t_right);
    take(R.t_first, fresh.t_first), take(R.t_last, fresh.t_last), take(R.t_pint, fresh.t_pint);
    take(R.t_pleaf, fresh.t_pleaf), take(R.t_newidx, fresh.t_newidx);
    R.topo_n = fresh.topo_n, R.topo_emit = fresh.topo_emit;
    R.topo_hash = th;
    R.view = fresh.view;
    R.camera = fresh.camera;
    R.build_ms = fresh.build_ms;
    R.bytes = fresh.bytes;
    R.n_nodes = fresh.n_nodes;
    R.lights_on = fresh.lights_on;
    R.max_fanout = fresh.max_fanout;
    R.geom_hash = gh;
    if (refit && fresh.topo_n > 0)
        ++R.refits;
    else
        ++R.rebuilds;
    return RTG_OK;
}

int rtg_scene_update(rtg_scene_t scene, const rtg_scene_desc *desc) { return scene_update(scene, desc, true); }

int rtg_scene_rebuild(rtg_scene_t scene, const rtg_scene_desc *desc) { return scene_update(scene, desc, false); }

int rtg_scene_stats(rtg_scene_t scene, rtg_scene_info *out) {
    if (scene == nullptr || out == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    const SceneRes &R = scene->res;
    out->bytes = R.bytes;
    out->build_ms = R.build_ms;
    out->n_nodes = R.n_nodes;
    out->n_prims = R.view.n_prims;
    out->rebuilds = R.rebuilds;
    out->refits = R.refits;
    out->shading_updates = R.shading_updates;
    return RTG_OK;
}

int rtg_scene_destroy(rtg_scene_t scene) {
    if (scene == nullptr) return RTG_OK;
    if (arena()) {
        std::lock_guard<std::mutex> l(arena()->lock);
        (void)hipStreamSynchronize(arena()->stream);
    }
    delete scene;
    return RTG_OK;
}

int rtg_scene_bytes(rtg_scene_t scene, uint64_t *bytes) {
    if (scene == nullptr || bytes == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    *bytes = scene->res.bytes;
    return RTG_OK;
}

int rtg_scene_build_ms(rtg_scene_t scene, double *ms) {
    if (scene == nullptr || ms == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    *ms = scene->res.build_ms;
    return RTG_OK;
}

// Grow the arena's pinned result staging to `bytes` (the caller holds A.lock;
// no copy into it is in flight: every batch synchronises before it returns).
static int ensure_pinned_out(Arena &A, size_t bytes) {
    if (A.pinned_out_bytes >= bytes) return RTG_OK;
    if (A.pinned_out) (void)hipHostFree(A.pinned_out);
    A.pinned_out = nullptr;
    A.pinned_out_bytes = 0;
    const size_t n = bytes + bytes / 4;
    RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&A.pinned_out), n, hipHostMallocDefault));
    A.pinned_out_bytes = n;
    return RTG_OK;
}

// Run copy(i) for i in [0, n) on up to 8 host threads (the calling thread
// included); one thread if no more can be started.
extern "C++" template <class F>
static void parallel_copies(int n, F &&copy) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < n; i = next++) copy(i);
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({8, n, (int)hw}));
    std::vector<std::thread> pool;
    try {
        for (int k = 1; k < nt; ++k) pool.emplace_back(work);
    } catch (...) {
        // fewer threads: the ones started and this one finish the work
    }
    work();
    for (std::thread &t : pool) t.join();
}

}  // extern "C"

namespace rtg {
void copy_out(const HostCopy *cp, int n) {
    constexpr size_t kPiece = size_t(1) << 20;
    std::vector<HostCopy> pieces;
    for (int i = 0; i < n; ++i)
        for (size_t off = 0; off < cp[i].bytes; off += kPiece)
            pieces.push_back(HostCopy{static_cast<uint8_t *>(cp[i].dst) + off, cp[i].src + off,
                                      std::min(kPiece, cp[i].bytes - off)});
    parallel_copies((int)pieces.size(), [&](int k) { std::memcpy(pieces[k].dst, pieces[k].src, pieces[k].bytes); });
}
}  // namespace rtg

extern "C" {

static int render_host(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb, int32_t *ids, float *tprim,
                       rtg_stats *stats, rtg_job_s *job) {
    if (scene == nullptr || rgb == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    int W, H;
    int st = resolve_size(scene->res, params, W, H);
    if (st != RTG_OK) return st;
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    const size_t npix = (size_t)W * H;
    if (ensure(A.out_rgb, npix * 3) || (ids && ensure(A.out_ids, npix * 8)) || (tprim && ensure(A.out_t, npix * 4)))
        return fail(RTG_E_NOMEM, "output buffers");
    // The arena's mapped cancel flag belongs to this job while it renders:
    // rtg_cancel sets it only for the arena's current job, under cancel_lock,
    // and it is cleared when the job ends.
    struct JobFlag {  // owns the flag for the job's render, on every return path
        Arena &A;
        rtg_job_s *job;
        bool seen = false;  // the flag was set when the job gave it up
        JobFlag(Arena &a, rtg_job_s *j) : A(a), job(j) {
            if (!job) return;
            std::lock_guard<std::mutex> c(A.cancel_lock);
            A.cancel_job = job;
            *A.cancel_host = job->cancel.load();
        }
        void release() {
            if (!job) return;
            std::lock_guard<std::mutex> c(A.cancel_lock);
            seen = seen || *A.cancel_host != 0;
            *A.cancel_host = 0;
            A.cancel_job = nullptr;
        }
        ~JobFlag() { release(); }
    } flag(A, job);
    // A cancellable frame renders in passes of at most 2^19 samples (whole
    // pixel-block rows): a cancel stops the pass in flight, and the passes
    // already done keep their pixels, as the reference's workers keep the
    // tiles they finished.
    const Cancel cancel{A.cancel_host, A.cancel_dev, size_t(1) << 19};
    RenderOut o{A.out_rgb.as<uint8_t>(), ids ? A.out_ids.as<int32_t>() : nullptr, tprim ? A.out_t.as<float>() : nullptr,
                A.counters.as<unsigned long long>(), A.counters.as<unsigned long long>() + 4};
    RTG_HIP(hipEventRecord(A.ev0, A.stream));
    st = render_device(scene->res, *params, W, H, o, A.stream, A, job ? &cancel : nullptr);
    RTG_HIP(hipEventRecord(A.ev1, A.stream));
    // Results through the arena's pinned staging: one device-to-host copy per
    // array, then host copies into the caller's (pageable) arrays.
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    const size_t b_rgb = npix * 3, b_ids = ids ? npix * 8 : 0, b_t = tprim ? npix * 4 : 0;
    const size_t o_ids = (b_rgb + 255) & ~size_t(255), o_t = (o_ids + b_ids + 255) & ~size_t(255);
    const int pst = ensure_pinned_out(A, o_t + b_t);
    if (pst == RTG_OK) {
        RTG_HIP(hipMemcpyAsync(A.pinned_out, o.rgb, b_rgb, hipMemcpyDeviceToHost, A.stream));
        if (ids) RTG_HIP(hipMemcpyAsync(A.pinned_out + o_ids, o.ids, b_ids, hipMemcpyDeviceToHost, A.stream));
        if (tprim) RTG_HIP(hipMemcpyAsync(A.pinned_out + o_t, o.tprim, b_t, hipMemcpyDeviceToHost, A.stream));
    }
    RTG_HIP(hipMemcpyAsync(cnt, o.counts, sizeof(cnt), hipMemcpyDeviceToHost, A.stream));  // counts + re-tests
    RTG_HIP(hipStreamSynchronize(A.stream));
    if (pst == RTG_OK) {
        const HostCopy cp[3] = {{rgb, A.pinned_out, b_rgb}, {ids, A.pinned_out + o_ids, b_ids},
                                {tprim, A.pinned_out + o_t, b_t}};
        copy_out(cp, 3);
    }
    flag.release();
    const bool cancelled = flag.seen;
    if (st != RTG_OK) return st;
    if (pst != RTG_OK) return pst;
    float ms = 0;
    RTG_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    if (stats) {
        stats->rays_primary = cnt[0];
        stats->rays_shadow = cnt[1];
        stats->rays_reflect = cnt[2];
        stats->rays_refract = cnt[3];
        stats->rays_total = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        stats->render_ms = ms;
        stats->build_ms = scene->res.build_ms;
        stats->retests_fp64 = cnt[4];
    }
    if (cancelled) return fail(RTG_E_CANCELLED, "render cancelled");
    return RTG_OK;
}

int rtg_render(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb, int32_t *ids, float *tprim,
               rtg_stats *stats) {
    return render_host(scene, params, rgb, ids, tprim, stats, nullptr);
}

int rtg_render_device(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb_dev, int32_t *ids_dev,
                      float *tprim_dev, uint64_t *stats_dev, void *stream) {
    if (scene == nullptr || rgb_dev == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    int W, H;
    int st = resolve_size(scene->res, params, W, H);
    if (st != RTG_OK) return st;
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : A.stream;
    unsigned long long *counts =
        stats_dev ? reinterpret_cast<unsigned long long *>(stats_dev) : A.counters.as<unsigned long long>();
    RenderOut o{rgb_dev, ids_dev, tprim_dev, counts, A.counters.as<unsigned long long>() + 4};
    return render_device(scene->res, *params, W, H, o, s, A, nullptr);
}

int rtg_render_async(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb, int32_t *ids, float *tprim,
                     rtg_job_t *job) {
    if (job == nullptr || params == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    rtg_job_s *j = new rtg_job_s();
    const rtg_params p = *params;
    j->th = std::thread([j, scene, p, rgb, ids, tprim]() {
        const int st = render_host(scene, &p, rgb, ids, tprim, &j->stats, j);
        j->status = st;
    });
    *job = j;
    return RTG_OK;
}

int rtg_poll(rtg_job_t job) {
    if (job == nullptr) return fail(RTG_E_INVALID, "job is NULL");
    return job->status.load();
}

int rtg_cancel(rtg_job_t job) {
    if (job == nullptr) return fail(RTG_E_INVALID, "job is NULL");
    job->cancel = 1;
    if (Arena *A = arena()) {
        std::lock_guard<std::mutex> c(A->cancel_lock);
        if (A->cancel_job == job) *A->cancel_host = 1;
    }
    return RTG_OK;
}

int rtg_job_wait(rtg_job_t job, rtg_stats *stats) {
    if (job == nullptr) return fail(RTG_E_INVALID, "job is NULL");
    if (job->th.joinable()) job->th.join();
    if (stats) *stats = job->stats;
    return job->status.load();
}

int rtg_job_free(rtg_job_t job) {
    if (job == nullptr) return RTG_OK;
    if (job->th.joinable()) job->th.join();
    delete job;
    return RTG_OK;
}

int rtg_kernel_times(rtg_kernel_time *out, int max, int *n) {
    if (n == nullptr || (out == nullptr && max > 0)) return fail(RTG_E_INVALID, "NULL argument");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    static const char *names[KT_COUNT] = {"k_extend",  "k_shade",      "k_occlude", "k_combine",
                                          "k_resolve", "k_megakernel", "k_local"};
    rtg_kernel_time acc[KT_COUNT];
    memset(acc, 0, sizeof(acc));
    int st = RTG_OK;
    for (const Arena::KernelRec &r : A.krecs) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.e1) != hipSuccess || hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess)
            st = fail(RTG_E_DEVICE, "kernel timing events");
        rtg_kernel_time &a = acc[r.kind];
        a.launches++;
        a.items += r.slot >= 0 ? A.kitems_host[r.slot] : r.items;
        a.ms += ms;
        A.kevents.push_back(r.e0);
        A.kevents.push_back(r.e1);
    }
    A.krecs.clear();
    A.kitems_next = 0;
    int k = 0;
    for (int i = 0; i < KT_COUNT; ++i) {
        if (acc[i].launches == 0) continue;
        if (k < max) {
            out[k] = acc[i];
            snprintf(out[k].name, sizeof(out[k].name), "%s", names[i]);
        }
        ++k;
    }
    *n = k;
    return st;
}

int rtg_render_batch(const rtg_request *reqs, int n, rtg_stats *stats) {
    if (n < 0 || (n > 0 && reqs == nullptr)) return fail(RTG_E_INVALID, "bad request list");
    if (n > 4096) return fail(RTG_E_LIMIT, "more than 4096 requests in a batch");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    if (stats) *stats = rtg_stats{};
    if (n == 0) return RTG_OK;
    Arena &A = *arena();
    std::vector<BatchItem> items(n);
    std::vector<size_t> off_rgb(n), off_ids(n);
    size_t rgb_bytes = 0, ids_bytes = 0;
    bool eager = false;
    for (int i = 0; i < n; ++i) {
        const rtg_request &q = reqs[i];
        if (q.scene == nullptr || q.rgb == nullptr) return fail(RTG_E_INVALID, "request without scene or rgb");
        int W, H;
        int st = resolve_size(q.scene->res, &q.params, W, H);
        if (st != RTG_OK) return st;
        if (q.params.mode != RTG_MODE_RAYTRACE && q.params.mode != RTG_MODE_REFRACTTEST &&
            q.params.mode != RTG_MODE_REFLECTTEST && q.params.mode != 0)
            return fail(RTG_E_INVALID, "batches render the ray-traced modes (debug modes: one frame per call)");
        eager = eager || (q.params.flags & RTG_FLAG_EAGER) != 0;
        items[i] = BatchItem{&q.scene->res, q.params, W, H, nullptr, nullptr, nullptr};
        off_rgb[i] = rgb_bytes;
        rgb_bytes += ((size_t)W * H * 3 + 255) & ~size_t(255);
        off_ids[i] = ids_bytes;
        if (q.ids) ids_bytes += (size_t)W * H * 8;
    }
    std::lock_guard<std::mutex> l(A.lock);
    A.ktimes = false;
    if (ensure(A.out_rgb, rgb_bytes) || (ids_bytes && ensure(A.out_ids, ids_bytes)))
        return fail(RTG_E_NOMEM, "batch output buffers");
    for (int i = 0; i < n; ++i) {
        items[i].rgb = A.out_rgb.as<uint8_t>() + off_rgb[i];
        if (reqs[i].ids) items[i].ids = reinterpret_cast<int32_t *>(A.out_ids.as<uint8_t>() + off_ids[i]);
    }
    hipStream_t s = A.stream;
    unsigned long long *counts = A.counters.as<unsigned long long>();
    k_fill_u8<<<1024, 256, 0, s>>>(A.out_rgb.as<uint8_t>(), rgb_bytes, 127);
    if (ids_bytes) k_fill_u8<<<1024, 256, 0, s>>>(A.out_ids.as<uint8_t>(), ids_bytes, 0xff);
    k_zero_u64<<<1, 64, 0, s>>>(counts, 5);
    RTG_HIP(hipEventRecord(A.ev0, s));
    // One wavefront level holds at most level_ray_cap rays: a request list
    // with more level-0 samples runs as consecutive batches, and a request
    // beyond the cap on its own as a frame in passes.
    int n_on = 0;
    std::vector<size_t> samples(n);
    for (int i = 0; i < n; ++i) {
        n_on = std::max(n_on, __builtin_popcount(items[i].scene->lights_on));
        samples[i] = frame_samples(items[i].W, items[i].H, items[i].params.spp, A.tune.tile_lg);
    }
    const size_t cap = level_ray_cap(n_on);
    int replayed = 0;
    int st = A.ws.order_before(s);
    for (int b = 0; st == RTG_OK && b < n;) {
        const int e = batch_group_end(samples.data(), n, b, cap);
        if (e == b + 1 && batch_bucket(samples[b]) > cap) {
            const RenderOut o{items[b].rgb, items[b].ids, nullptr, counts, counts + 4};
            st = launch_wavefront(*items[b].scene, items[b].params, items[b].W, items[b].H, o, s, A.ws, A, nullptr);
        } else {
            int rep = 0;
            st = launch_batch(items.data() + b, e - b, counts, counts + 4, s, A.ws, A, eager ? 0 : 1, &rep);
            replayed |= rep;
        }
        b = e;
    }
    if (st == RTG_OK) st = A.ws.mark_done(s);
    RTG_HIP(hipEventRecord(A.ev1, s));
    // Results: one copy of the whole output region into pinned memory, then a
    // scatter into the callers' (pageable) arrays on host threads, instead of
    // one staged pageable copy per request.
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    if (st == RTG_OK) st = ensure_pinned_out(A, rgb_bytes + ids_bytes);
    if (st == RTG_OK) {
        RTG_HIP(hipMemcpyAsync(A.pinned_out, A.out_rgb.p, rgb_bytes, hipMemcpyDeviceToHost, s));
        if (ids_bytes)
            RTG_HIP(hipMemcpyAsync(A.pinned_out + rgb_bytes, A.out_ids.p, ids_bytes, hipMemcpyDeviceToHost, s));
        RTG_HIP(hipMemcpyAsync(cnt, counts, sizeof(cnt), hipMemcpyDeviceToHost, s));
    }
    RTG_HIP(hipStreamSynchronize(s));
    if (st != RTG_OK) return st;
    std::vector<HostCopy> cp;
    for (int i = 0; i < n; ++i) {
        const size_t npix = (size_t)items[i].W * items[i].H;
        cp.push_back(HostCopy{reqs[i].rgb, A.pinned_out + off_rgb[i], npix * 3});
        if (reqs[i].ids) cp.push_back(HostCopy{reqs[i].ids, A.pinned_out + rgb_bytes + off_ids[i], npix * 8});
    }
    copy_out(cp.data(), (int)cp.size());
    float ms = 0;
    RTG_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    if (stats) {
        stats->rays_primary = cnt[0];
        stats->rays_shadow = cnt[1];
        stats->rays_reflect = cnt[2];
        stats->rays_refract = cnt[3];
        stats->rays_total = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        stats->render_ms = ms;
        stats->graph_replays = (uint64_t)replayed;
        stats->retests_fp64 = cnt[4];
    }
    return RTG_OK;
}

// ---- launch tuning ------------------------------------------------------

static int tuning_from(const rtg_tuning &t, Tuning &out) {
    const int lg = t.tile_w == 4 ? 2 : t.tile_w == 8 ? 3 : t.tile_w == 16 ? 4 : t.tile_w == 32 ? 5 : -1;
    const int g = t.grab == 1 ? 0 : t.grab == 2 ? 1 : t.grab == 4 ? 2 : -1;
    if (lg < 0) return fail(RTG_E_INVALID, "tuning: tile_w must be 4, 8, 16 or 32");
    if (g < 0) return fail(RTG_E_INVALID, "tuning: grab must be 1, 2 or 4");
    if (t.grid_pct < 25 || t.grid_pct > 100) return fail(RTG_E_INVALID, "tuning: grid_pct must be 25..100");
    out.tile_lg = lg, out.grab_lg = g, out.grid_pct = t.grid_pct;
    return RTG_OK;
}

static rtg_tuning tuning_to(const Tuning &t) { return rtg_tuning{1 << t.tile_lg, 1 << t.grab_lg, t.grid_pct, 0}; }

int rtg_get_tuning(rtg_tuning *out) {
    if (out == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    *out = tuning_to(A.tune);
    return RTG_OK;
}

int rtg_set_tuning(const rtg_tuning *t) {
    if (t == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    Tuning v;
    const int st = tuning_from(*t, v);
    if (st != RTG_OK) return st;
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    A.tune = v;
    return RTG_OK;
}

int rtg_autotune(const rtg_request *reqs, int n, int reps, rtg_tuning *best, rtg_tune_trial *trials, int max_trials,
                 int *n_trials) {
    if (n < 1 || reqs == nullptr) return fail(RTG_E_INVALID, "autotune needs at least one request");
    if (n > 4096) return fail(RTG_E_LIMIT, "more than 4096 requests in a batch");
    if (reps < 1 || reps > 100) return fail(RTG_E_INVALID, "reps must be 1..100");
    if (max_trials < 0 || (max_trials > 0 && trials == nullptr)) return fail(RTG_E_INVALID, "bad trial buffer");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    std::vector<BatchItem> items(n);
    std::vector<size_t> off_rgb(n);
    size_t rgb_bytes = 0;
    for (int i = 0; i < n; ++i) {
        const rtg_request &q = reqs[i];
        if (q.scene == nullptr) return fail(RTG_E_INVALID, "request without scene");
        int W, H;
        const int st = resolve_size(q.scene->res, &q.params, W, H);
        if (st != RTG_OK) return st;
        if (q.params.mode != RTG_MODE_RAYTRACE && q.params.mode != RTG_MODE_REFRACTTEST &&
            q.params.mode != RTG_MODE_REFLECTTEST && q.params.mode != 0)
            return fail(RTG_E_INVALID, "the wavefront engine renders the ray-traced modes");
        items[i] = BatchItem{&q.scene->res, q.params, W, H, nullptr, nullptr, nullptr};
        off_rgb[i] = rgb_bytes;
        rgb_bytes += ((size_t)W * H * 3 + 255) & ~size_t(255);
    }
    std::lock_guard<std::mutex> l(A.lock);
    A.ktimes = false;
    if (ensure(A.out_rgb, rgb_bytes)) return fail(RTG_E_NOMEM, "autotune output buffers");
    for (int i = 0; i < n; ++i) items[i].rgb = A.out_rgb.as<uint8_t>() + off_rgb[i];
    hipStream_t s = A.stream;
    unsigned long long *counts = A.counters.as<unsigned long long>();
    auto run = [&]() -> int {
        k_zero_u64<<<1, 64, 0, s>>>(counts, 5);
        int st = A.ws.order_before(s);
        if (st != RTG_OK) return st;
        if (n == 1) {
            const RenderOut o{items[0].rgb, nullptr, nullptr, counts, counts + 4};
            st = launch_wavefront(*items[0].scene, items[0].params, items[0].W, items[0].H, o, s, A.ws, A, nullptr);
        } else {
            st = launch_batch(items.data(), n, counts, counts + 4, s, A.ws, A, 0, nullptr);
        }
        const int st2 = A.ws.mark_done(s);
        return st != RTG_OK ? st : st2;
    };
    // median device time of `reps` runs after one warmup run
    auto time_of = [&](const Tuning &t, double &ms) -> int {
        A.tune = t;
        int st = run();
        if (st != RTG_OK) return st;
        std::vector<double> v;
        for (int r = 0; r < reps; ++r) {
            RTG_HIP(hipEventRecord(A.ev0, s));
            if ((st = run()) != RTG_OK) return st;
            RTG_HIP(hipEventRecord(A.ev1, s));
            RTG_HIP(hipEventSynchronize(A.ev1));
            float e = 0.0f;
            RTG_HIP(hipEventElapsedTime(&e, A.ev0, A.ev1));
            v.push_back(e);
        }
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        return RTG_OK;
    };
    const Tuning start = A.tune;
    std::vector<std::pair<uint64_t, double>> seen;  // tuning key -> median ms
    int count = 0;
    auto trial = [&](const Tuning &t, double &ms) -> int {
        for (const auto &e : seen)
            if (e.first == t.key()) {
                ms = e.second;
                return RTG_OK;
            }
        const int st = time_of(t, ms);
        if (st != RTG_OK) return st;
        seen.emplace_back(t.key(), ms);
        if (count < max_trials) trials[count] = rtg_tune_trial{tuning_to(t), ms};
        ++count;
        return RTG_OK;
    };
    // coordinate descent: tile shape, then queue grab, then grid; repeated
    // while a pass changes anything (at most three passes)
    static const int tiles[] = {2, 3, 4, 5}, grabs[] = {0, 1, 2}, grids[] = {50, 75, 100};
    Tuning cur = start;
    double cur_ms = 0.0;
    int st = trial(cur, cur_ms);
    for (int pass = 0; st == RTG_OK && pass < 3; ++pass) {
        bool changed = false;
        for (int knob = 0; knob < 3 && st == RTG_OK; ++knob) {
            const int *vals = knob == 0 ? tiles : knob == 1 ? grabs : grids;
            const int nv = knob == 0 ? 4 : 3;
            for (int k = 0; k < nv && st == RTG_OK; ++k) {
                Tuning c = cur;
                (knob == 0 ? c.tile_lg : knob == 1 ? c.grab_lg : c.grid_pct) = vals[k];
                double ms = 0.0;
                st = trial(c, ms);
                if (st == RTG_OK && ms < cur_ms) {
                    changed = changed || c.key() != cur.key();
                    cur = c, cur_ms = ms;
                }
            }
        }
        if (!changed) break;
    }
    RTG_HIP(hipStreamSynchronize(s));
    A.tune = st == RTG_OK ? cur : start;
    if (st != RTG_OK) return st;
    if (best) *best = tuning_to(cur);
    if (n_trials) *n_trials = count;
    return RTG_OK;
}

}  // extern "C"
