// rtg_serve.hip — the request loop: open-loop arrivals, admission by HBM
// budget, continuous batching over concurrent slots (rtg_serve).
//
// Batched mode.  A slot is a stream with its own workspace (queues, spill,
// tables, captured graphs), device output buffer and pinned host staging.
// The loop admits every request whose arrival time has passed; whenever a
// slot is free it forms one batch of waiting requests (first come first
// served, up to max_batch and the HBM budget of the batch's queues) and
// enqueues it as one wavefront on that slot, followed by the copy of the
// batch's RGB to pinned memory and an event.  Several slots run at once, so
// a batch admitted while others are on the GPU overlaps them on the device.
// A completed slot's RGB is copied into each request's buffer and the
// request is done.  A batch shape of at least kGraphMinBatch requests
// launched `graph_after` times is captured into a hipGraph and replayed from
// then on.
//
// Plain mode.  One slot, one request at a time, each rendered as its own
// frame (the single-frame wavefront), in arrival order: the plain renderer
// the loop is compared with.  (A batched-mode batch of one request takes the
// same single-frame path.)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "rt_prep.hpp"
#include "rtg_internal.hpp"

namespace rtg {
namespace {

__global__ void k_fill_bytes(uint8_t *p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_counts(unsigned long long *p) {
    if (threadIdx.x < 5) p[threadIdx.x] = 0;  // 4 ray counts + FP64 re-tests
}

using Clock = std::chrono::steady_clock;

constexpr int kGraphMinBatch = 8;  // smallest batch the loop captures into a hipGraph

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

int slot_init(ServeSlot &s) {
    if (s.stream == nullptr) RTG_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    if (s.done == nullptr) RTG_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    if (s.counts_host == nullptr)
        RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.counts_host), 4 * sizeof(unsigned long long),
                              hipHostMallocDefault));
    if (ensure(s.counts, 5 * sizeof(unsigned long long))) return fail(RTG_E_NOMEM, "serve counters");
    return RTG_OK;
}

int slot_reserve(ServeSlot &s, size_t bytes) {
    if (ensure(s.out, bytes)) return fail(RTG_E_NOMEM, "serve outputs");
    if (s.pinned_bytes < bytes) {
        if (s.pinned) (void)hipHostFree(s.pinned);
        s.pinned = nullptr;
        s.pinned_bytes = 0;
        const size_t n = bytes + bytes / 4;
        RTG_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.pinned), n, hipHostMallocDefault));
        s.pinned_bytes = n;
    }
    return RTG_OK;
}

struct ReqInfo {
    int W, H;
    size_t samples;
    int fanout, n_on;
};

}  // namespace

ServeSlot::~ServeSlot() {
    if (stream) (void)hipStreamSynchronize(stream);
    ws.release();
    if (pinned) (void)hipHostFree(pinned);
    if (counts_host) (void)hipHostFree(counts_host);
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
}

// When the loop is saturated (another slot's batch in flight and requests
// still waiting), a batch goes on its slot's stream alone: the slots already
// keep the GPU busy side by side, and a side stream per busy slot would put
// more streams than the process has hardware queues (4), so that streams of
// different slots share a queue and wait for each other (capacity at 6000
// req/s offered, 2 slots: 4215 req/s with side streams, 4743 without).
// Below saturation a batch keeps the side stream, whose overlap shortens its
// latency (2000 req/s: batched p99 12.4 ms with it, 19.4 ms without).
static rtg_params stream_mode(rtg_params p, bool side_stream) {
    if (!side_stream) p.flags |= RTG_FLAG_ONE_STREAM;
    return p;
}

// Enqueue `ids` as one batch on slot `s` (asynchronous; `s.done` marks the
// end of the RGB copy to pinned memory).
static int slot_launch(ServeSlot &s, rtg_serve_request *reqs, const std::vector<ReqInfo> &info,
                       const std::vector<int> &ids, bool plain, int capture_at, Arena &A, bool side_stream) {
    size_t bytes = 0;
    s.offs.resize(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
        s.offs[k] = bytes;
        bytes += ((size_t)info[ids[k]].W * info[ids[k]].H * 3 + 255) & ~size_t(255);
    }
    int st = slot_reserve(s, bytes);
    if (st != RTG_OK) return st;
    k_fill_bytes<<<1024, 256, 0, s.stream>>>(s.out.as<uint8_t>(), bytes, 127);
    k_zero_counts<<<1, 64, 0, s.stream>>>(s.counts.as<unsigned long long>());
    if (plain || ids.size() == 1) {  // a batch of one takes the single-frame path (uniform kernel arguments)
        const rtg_serve_request &q = reqs[ids[0]];
        const RenderOut o{s.out.as<uint8_t>(), nullptr, nullptr, s.counts.as<unsigned long long>(),
                          s.counts.as<unsigned long long>() + 4};
        st = launch_wavefront(q.scene->res, stream_mode(q.params, side_stream), info[ids[0]].W, info[ids[0]].H, o, s.stream, s.ws, A,
                              nullptr);
    } else {
        std::vector<BatchItem> items(ids.size());
        for (size_t k = 0; k < ids.size(); ++k) {
            const rtg_serve_request &q = reqs[ids[k]];
            items[k] = BatchItem{&q.scene->res, stream_mode(q.params, side_stream), info[ids[k]].W, info[ids[k]].H,
                                 s.out.as<uint8_t>() + s.offs[k], nullptr, nullptr};
        }
        // A small batch launches eagerly: capturing one of its many shapes costs
        // milliseconds inside some request's latency for little launch time saved.
        int replayed = 0;
        const int cap = (int)items.size() >= kGraphMinBatch ? capture_at : 0;
        st = launch_batch(items.data(), (int)items.size(), s.counts.as<unsigned long long>(),
                          s.counts.as<unsigned long long>() + 4, s.stream, s.ws, A, cap, &replayed);
    }
    if (st != RTG_OK) return st;
    RTG_HIP(hipMemcpyAsync(s.pinned, s.out.p, bytes, hipMemcpyDeviceToHost, s.stream));
    RTG_HIP(hipMemcpyAsync(s.counts_host, s.counts.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           s.stream));
    RTG_HIP(hipEventRecord(s.done, s.stream));
    s.reqs = ids;
    s.busy = true;
    return RTG_OK;
}

int serve_run(rtg_serve_request *reqs, int n, const rtg_serve_opts &o, rtg_serve_stats &out, Arena &A) {
    const bool plain = o.mode == RTG_SERVE_PLAIN;
    const int nslots = plain ? 1 : (o.slots > 0 ? std::min(o.slots, 8) : 2);
    const int max_batch = plain ? 1 : (o.max_batch > 0 ? std::min(o.max_batch, 4096) : 256);
    const size_t budget = o.hbm_budget > 0 ? (size_t)o.hbm_budget : (size_t(4) << 30);
    const int capture_at = o.graph_after < 0 ? 0 : (o.graph_after == 0 ? 3 : o.graph_after);
    std::vector<ReqInfo> info(n);
    for (int i = 0; i < n; ++i) {
        rtg_serve_request &q = reqs[i];
        q.status = RTG_E_PENDING;
        q.done_s = 0.0;
        q.batch = -1;
        if (q.scene == nullptr || q.rgb == nullptr) return fail(RTG_E_INVALID, "request without scene or rgb");
        int st = resolve_size(q.scene->res, &q.params, info[i].W, info[i].H);
        if (st != RTG_OK) return st;
        if (q.params.mode != RTG_MODE_RAYTRACE && q.params.mode != RTG_MODE_REFRACTTEST &&
            q.params.mode != RTG_MODE_REFLECTTEST && q.params.mode != 0)
            return fail(RTG_E_INVALID, "the serving loop renders the ray-traced modes");
        info[i].samples = frame_samples(info[i].W, info[i].H, q.params.spp, A.tune.tile_lg);
        info[i].fanout = q.scene->res.max_fanout;
        info[i].n_on = __builtin_popcount(q.scene->res.lights_on);
    }
    while ((int)A.serve_slots.size() < nslots) A.serve_slots.emplace_back(new ServeSlot());
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    // Size every slot, before the clock starts, for the largest single
    // request and the largest batch the loop can form, so that no request
    // waits for a buffer to grow (a reallocation waits for the whole device).
    // A batch is a run of at most max_batch consecutive requests in arrival
    // order: the largest such window bounds its samples, RGB bytes and table
    // entries, and its queue memory stays within the budget.
    int maxW = 1, maxH = 1, max_spp = 1, max_lvl = 0, max_fan = 0, max_on = 0;
    size_t max_rgb = 0;
    auto rgb_bytes = [&](int i) { return ((size_t)info[i].W * info[i].H * 3 + 255) & ~size_t(255); };
    for (int i = 0; i < n; ++i) {
        maxW = std::max(maxW, info[i].W), maxH = std::max(maxH, info[i].H);
        max_spp = std::max(max_spp, reqs[i].params.spp > 0 ? (int)reqs[i].params.spp : 1);
        max_lvl = std::max(max_lvl, (int)reqs[i].params.max_level);
        max_fan = std::max(max_fan, info[i].fanout), max_on = std::max(max_on, info[i].n_on);
        max_rgb = std::max(max_rgb, rgb_bytes(i));
    }
    size_t win_samples = 0, win_rgb = max_rgb, win_tab = 0;
    int n_scenes = 0;
    if (max_batch > 1) {
        size_t sum_s = 0, sum_r = 0, sum_t = 0;
        for (int j = 0; j < n; ++j) {
            const int i = order[j];
            sum_s += info[i].samples, sum_r += rgb_bytes(i), sum_t += (size_t)info[i].W + info[i].H;
            if (j >= max_batch) {
                const int o = order[j - max_batch];
                sum_s -= info[o].samples, sum_r -= rgb_bytes(o), sum_t -= (size_t)info[o].W + info[o].H;
            }
            win_samples = std::max(win_samples, sum_s), win_rgb = std::max(win_rgb, sum_r);
            win_tab = std::max(win_tab, sum_t);
        }
        // the largest level-0 length whose queues fit the budget
        size_t lo = 0, hi = win_samples;
        while (lo < hi) {
            const size_t mid = lo + (hi - lo + 1) / 2;
            const size_t need = batch_queue_bytes(batch_bucket(mid), max_lvl, max_fan, max_on, A.no_sync_budget);
            if (need != 0 && need <= budget)
                lo = mid;
            else
                hi = mid - 1;
        }
        win_samples = lo;
        std::vector<const rtg_scene_s *> seen;
        for (int i = 0; i < n; ++i)
            if (std::find(seen.begin(), seen.end(), reqs[i].scene) == seen.end()) seen.push_back(reqs[i].scene);
        n_scenes = (int)std::min<size_t>(seen.size(), (size_t)max_batch);
    }
    for (int k = 0; k < nslots; ++k) {
        ServeSlot &s = *A.serve_slots[k];
        int st = slot_init(s);
        if (st == RTG_OK && n > 0) st = slot_reserve(s, win_rgb);
        if (st == RTG_OK && n > 0) st = reserve_frame(s.ws, maxW, maxH, max_spp, max_lvl, max_fan, max_on, A);
        if (st == RTG_OK && win_samples > 0)
            st = reserve_batch(s.ws, win_samples, std::min(max_batch, n), n_scenes, win_tab, max_lvl, max_fan, max_on,
                               s.stream, A);
        if (st != RTG_OK) return st;
    }

    std::vector<int> waiting;  // arrived, not yet admitted (arrival order)
    size_t head = 0;           // first waiting entry not yet admitted
    int next = 0, completed = 0, batches = 0, in_flight = 0, max_in_flight = 0, st = RTG_OK;
    uint64_t rays = 0;
    const Clock::time_point t0 = Clock::now();
    while (completed < n) {
        bool progress = false;
        double now = since(t0);
        while (next < n && reqs[order[next]].arrival_s <= now) waiting.push_back(order[next++]);
        // completions
        for (int k = 0; k < nslots; ++k) {
            ServeSlot &s = *A.serve_slots[k];
            if (!s.busy) continue;
            const hipError_t e = hipEventQuery(s.done);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) {
                st = fail(RTG_E_DEVICE, std::string("serve batch: ") + hipGetErrorString(e));
                break;
            }
            std::vector<HostCopy> cp(s.reqs.size());
            for (size_t j = 0; j < s.reqs.size(); ++j)
                cp[j] = HostCopy{reqs[s.reqs[j]].rgb, s.pinned + s.offs[j],
                                 (size_t)info[s.reqs[j]].W * info[s.reqs[j]].H * 3};
            copy_out(cp.data(), (int)cp.size());
            const double t = since(t0);
            for (int id : s.reqs) {
                reqs[id].done_s = t;
                reqs[id].status = RTG_OK;
            }
            rays += s.counts_host[0] + s.counts_host[1] + s.counts_host[2] + s.counts_host[3];
            completed += (int)s.reqs.size();
            in_flight -= (int)s.reqs.size();
            s.busy = false;
            progress = true;
        }
        if (st != RTG_OK) break;
        // admissions: fill free slots with waiting requests
        for (int k = 0; k < nslots && head < waiting.size(); ++k) {
            ServeSlot &s = *A.serve_slots[k];
            if (s.busy) continue;
            std::vector<int> ids;
            size_t samples = 0;
            int lvl = 0, fan = 0, on = 0;
            while (head < waiting.size() && (int)ids.size() < max_batch) {
                const int id = waiting[head];
                const ReqInfo &r = info[id];
                const int l2 = std::max(lvl, (int)reqs[id].params.max_level), f2 = std::max(fan, r.fanout),
                          o2 = std::max(on, r.n_on);
                const size_t need = batch_queue_bytes(batch_bucket(samples + r.samples), l2, f2, o2, A.no_sync_budget);
                if (!ids.empty() && (need == 0 || need > budget)) break;  // the next batch takes it
                ids.push_back(id);
                samples += r.samples, lvl = l2, fan = f2, on = o2;
                ++head;
            }
            bool others = false;  // another slot's batch in flight
            for (int o = 0; o < nslots; ++o) others = others || (o != k && A.serve_slots[o]->busy);
            const bool saturated = others && head < waiting.size();
            st = slot_launch(s, reqs, info, ids, plain, capture_at, A, !saturated);
            if (st != RTG_OK) break;
            for (int id : ids) reqs[id].batch = batches;
            ++batches;
            in_flight += (int)ids.size();
            max_in_flight = std::max(max_in_flight, in_flight);
            progress = true;
        }
        if (st != RTG_OK) break;
        if (!progress) {
            // nothing to do until a slot completes or the next request arrives
            double wait = 50e-6;
            if (head >= waiting.size() && next < n && in_flight == 0)
                wait = std::max(0.0, std::min(reqs[order[next]].arrival_s - since(t0), 2e-3));
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }
    }
    if (st != RTG_OK) {
        for (int k = 0; k < nslots; ++k) (void)hipStreamSynchronize(A.serve_slots[k]->stream);
        for (auto &s : A.serve_slots) s->busy = false;
        return st;
    }
    std::vector<double> lat(n);
    double first = 1e300, last = 0.0, sum = 0.0, mx = 0.0;
    for (int i = 0; i < n; ++i) {
        lat[i] = (reqs[i].done_s - reqs[i].arrival_s) * 1e3;
        first = std::min(first, reqs[i].arrival_s);
        last = std::max(last, reqs[i].done_s);
        sum += lat[i];
        mx = std::max(mx, lat[i]);
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) { return n ? lat[std::min(n - 1, (int)std::ceil(p * n) - 1)] : 0.0; };
    out = rtg_serve_stats{};
    out.wall_s = n ? last - first : 0.0;
    out.p50_ms = pct(0.50);
    out.p99_ms = pct(0.99);
    out.mean_ms = n ? sum / n : 0.0;
    out.max_ms = mx;
    out.rays_total = rays;
    out.mrays_per_s = out.wall_s > 0 ? rays / out.wall_s / 1e6 : 0.0;
    out.req_per_s = out.wall_s > 0 ? n / out.wall_s : 0.0;
    out.batches = batches;
    out.max_in_flight = max_in_flight;
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" int rtg_serve(rtg_serve_request *reqs, int n, const rtg_serve_opts *opts, rtg_serve_stats *stats) {
    if (n < 0 || (n > 0 && reqs == nullptr)) return fail(RTG_E_INVALID, "bad request list");
    if (!arena()) return fail(RTG_E_STATE, "call rtg_init first");
    rtg_serve_opts o{};
    if (opts) o = *opts;
    if (o.mode != RTG_SERVE_BATCHED && o.mode != RTG_SERVE_PLAIN) return fail(RTG_E_INVALID, "unknown serve mode");
    Arena &A = *arena();
    std::lock_guard<std::mutex> l(A.lock);
    A.ktimes = false;
    rtg_serve_stats s{};
    const int st = serve_run(reqs, n, o, s, A);
    if (stats) *stats = s;
    return st;
}
