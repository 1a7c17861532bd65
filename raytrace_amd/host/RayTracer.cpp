// RayTracer.cpp — the drop-in render call, rewired onto librtg.
//
// start() keeps the reference's protocol (RayTracer.cpp:614-696): it returns at
// once, `isFinish` becomes true when the frame is complete, `useTime` holds the
// seconds it took, and stop() cancels cooperatively, leaving unrendered tiles
// at 127.  The output buffer is the reference's fixed 2048x2048x3 bytes with
// row stride 3*width, written in whole 64x64 tiles.
#include "RayTracer.h"
#include "flatten.h"

#include <chrono>

RayTracer::RayTracer(Scene &s) : scene(&s), flat(new SceneFlat()) {
    output = new uint8_t[2048 * 2048 * 3];
    memset(output, 127, 2048 * 2048 * 3);
}

RayTracer::~RayTracer() {
    stop();
    joinMonitor();
    if (devScene != nullptr) rtg_scene_destroy(devScene);
    delete[] output;
}

void RayTracer::joinMonitor() {
    if (monitor.joinable()) monitor.join();
}

void RayTracer::start(const uint8_t type, const int8_t tnum) {
    (void)tnum;
    joinMonitor();
    isFinish = false;
    width = scene->cam.width;
    height = scene->cam.height;
    memset(output, 127, 2048 * 2048 * 3);
    auto fail = [this](int st) {
        lastStatus = st;
        useTime = 0.0;
        isFinish = true;
    };
    if (width <= 0 || height <= 0 || width > 2048 || height > 2048) return fail(RTG_E_LIMIT);
    int st = rtg_init(0);
    if (st != RTG_OK) return fail(st);
    SceneFlattener::flatten(*scene, *flat);
    // The scene stays resident between frames; it is re-uploaded (and its BVH
    // rebuilt) only when the flattened content changed.
    const uint64_t fp = rtg_desc_fingerprint(flat->desc);
    if (devScene == nullptr || fp != devFingerprint) {
        st = devScene == nullptr ? rtg_scene_create(&flat->desc, &devScene) : rtg_scene_update(devScene, &flat->desc);
        if (st != RTG_OK) return fail(st);
        devFingerprint = fp;
        ++uploads;
    }
    rtg_params p{};
    p.mode = type;
    p.max_level = maxLevel;
    p.width = width;
    p.height = height;
    p.spp = 1;
    p.tile_compat = 1;
    p.engine = engine;
    rtg_job_t j = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    st = rtg_render_async(devScene, &p, output, nullptr, nullptr, &j);
    if (st != RTG_OK) return fail(st);
    {
        std::lock_guard<std::mutex> g(jobLock);
        job = j;
    }
    monitor = std::thread([this, t0, j]() {
        rtg_stats s{};
        int rc = rtg_job_wait(j, &s);
        {
            std::lock_guard<std::mutex> g(jobLock);
            job = nullptr;
        }
        rtg_job_free(j);
        lastStats = s;
        lastStatus = rc;
        useTime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        isFinish = true;
    });
}

void RayTracer::stop() {
    std::lock_guard<std::mutex> g(jobLock);
    if (job != nullptr) rtg_cancel(job);
}
