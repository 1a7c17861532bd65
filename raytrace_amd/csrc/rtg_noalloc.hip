// rtg_noalloc.hip — rtg_render_device_noalloc: one frame into the caller's
// device outputs on the caller's stream, with no device allocation.
//
// For callers that cannot allocate device memory while they render (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// It renders with the megakernel, whose scratch is bounded by the persistent
// grid rather than by the frame: the ray-generation tables (W + H floats) and
// the traversal stacks' spill area (80 ints per thread of the grid).  The
// wavefront engine's queues grow with the frame (about 1 KB per sample at
// depth 4, 2.2 GB for a 1080p frame), too much for a fixed array.
//
// That scratch is `scratch`, a file-scope device array of kScratchBytes: each
// GPU holds its own copy for the life of the process.  Every call on a GPU
// shares it and it keeps the last call's contents, so calls that overlap on
// one GPU (two streams, two threads) race; the counters a call reads are
// cleared by a kernel at its start.
#include "rt_devstack.hpp"
#include "rtg_internal.hpp"

namespace rtg {

namespace {

// Layout of the scratch: [0, 256) ray counts when the caller passes none;
// [256, kTablesEnd) ray-generation tables; [kTablesEnd, end) spill area.
constexpr size_t kCountsBytes = 256;
constexpr size_t kTablesEnd = kCountsBytes + size_t(2) * RTG_MAX_DIM * sizeof(float);
constexpr size_t kSpillPerBlock = size_t(kBlockThreads) * (kStackMax - kLdsStack) * sizeof(int);
constexpr int kMaxBlocks = 2048;  // 8 per CU on a 256-CU MI355X
constexpr size_t kScratchBytes = kTablesEnd + kMaxBlocks * kSpillPerBlock;  // 80 MiB + 128.25 KiB

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char *g_ptr[kMaxDevices] = {};  // the device address of `scratch` on each device

__global__ void k_fill_bytes(uint8_t *p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_counts(unsigned long long *p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Called by rtg_init with `device` current: the host name of `scratch` is a
// host-side stand-in, not a device address, so the address comes from the
// runtime, once per device.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(RTG_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_lock);
    if (g_ptr[device] != nullptr) return RTG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        return fail(RTG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<unsigned char *>(p);
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" int rtg_render_device_noalloc(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb_dev,
                                         int32_t *ids_dev, float *tprim_dev, uint64_t *stats_dev, void *stream) {
    if (scene == nullptr || rgb_dev == nullptr || params == nullptr) return fail(RTG_E_INVALID, "NULL argument");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    Arena &A = *arena();
    int W, H;
    int st = resolve_size(scene->res, params, W, H);  // W, H <= RTG_MAX_DIM: the tables fit their region
    if (st != RTG_OK) return st;
    if ((size_t)(W + H) * sizeof(float) > kTablesEnd - kCountsBytes)
        return fail(RTG_E_LIMIT, "frame larger than the no-allocation scratch's ray-generation tables");
    const rtg_params &p = *params;
    if (!(p.engine == RTG_ENGINE_AUTO || p.engine == RTG_ENGINE_MEGAKERNEL))
        return fail(RTG_E_INVALID, "rtg_render_device_noalloc renders with the megakernel (engine AUTO or MEGAKERNEL)");
    if (p.flags & (RTG_FLAG_FP64_RETEST | RTG_FLAG_KERNEL_TIMES))
        return fail(RTG_E_INVALID, "rtg_render_device_noalloc: no FP64 re-test (wavefront only) or kernel timing");
    if (!(p.mode == 0 || (p.mode >= RTG_MODE_CHECK && p.mode <= RTG_MODE_REFRACTTEST) || p.mode == RTG_MODE_RAYTRACE))
        return fail(RTG_E_INVALID, "unknown render mode");
    unsigned char *base = g_ptr[A.device];
    if (base == nullptr) return fail(RTG_E_STATE, "no-allocation scratch not set up on this device");
    std::lock_guard<std::mutex> l(A.lock);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : A.stream;
    MegaScratch ms{};
    ms.tables = reinterpret_cast<float *>(base + kCountsBytes);
    ms.spill = reinterpret_cast<int *>(base + kTablesEnd);
    ms.blocks = std::min(kMaxBlocks, persistent_blocks(A));
    ms.dp = tan((double)scene->res.camera.fovy * 3.1415926535897932384 / 360) / (H / 2);
    unsigned long long *counts = stats_dev ? reinterpret_cast<unsigned long long *>(stats_dev)
                                           : reinterpret_cast<unsigned long long *>(base);
    const size_t npix = (size_t)W * H;
    k_fill_bytes<<<1024, 256, 0, s>>>(rgb_dev, npix * 3, 127);
    if (ids_dev) k_fill_bytes<<<1024, 256, 0, s>>>(reinterpret_cast<uint8_t *>(ids_dev), npix * 8, 0xff);
    k_zero_counts<<<1, 64, 0, s>>>(counts);
    RTG_HIP(hipGetLastError());
    if ((st = write_raygen_tables(ms.tables, W, H, ms.dp, s)) != RTG_OK) return st;
    const bool kt = A.ktimes;
    A.ktimes = false;
    const RenderOut o{rgb_dev, ids_dev, tprim_dev, counts, nullptr};
    st = enqueue_megakernel(scene->res, p, W, H, o, s, A, nullptr, ms);
    A.ktimes = kt;
    return st;
}
