// RayTracer.h — drop-in renderer class (reference RayTracer.h:16-56).
//
// Same public surface: RayTracer(Scene&), start(type, tnum), stop(), output,
// isFinish, useTime, maxLevel, width, height.  start() flattens the scene,
// keeps it resident on the GPU while it is unchanged, and renders through the
// librtg C ABI on a helper thread; isFinish flips when the frame is done, as
// the reference's monitor thread does (RayTracer.cpp:674-695).
#pragma once
#include "Scene.h"
#include "../../include/rtg.h"

#include <mutex>

#define MY_MODEL_CHECK 0x1
#define MY_MODEL_DEPTHTEST 0x2
#define MY_MODEL_NORMALTEST 0x3
#define MY_MODEL_TEXTURETEST 0x4
#define MY_MODEL_MATERIALTEST 0x5
#define MY_MODEL_SHADOWTEST 0x6
#define MY_MODEL_REFLECTTEST 0x7
#define MY_MODEL_REFRACTTEST 0x8
#define MY_MODEL_RAYTRACE 0x80

struct SceneFlat;

class RayTracer {
private:
    Scene *scene;
    std::thread monitor;
    std::mutex jobLock;
    rtg_job_t job = nullptr;
    rtg_scene_t devScene = nullptr;
    uint64_t devFingerprint = 0;
    std::unique_ptr<SceneFlat> flat;
    void joinMonitor();
public:
    GLuint texID = 0;
    uint8_t *output;
    volatile bool isFinish = true;
    volatile double useTime = 0.0;
    uint8_t maxLevel = 1;
    int width = 0, height = 0;
    // Extensions: engine choice and the last frame's status and ray counts.
    int engine = RTG_ENGINE_AUTO;
    volatile int lastStatus = RTG_OK;
    rtg_stats lastStats{};
    uint64_t uploads = 0;  // scene uploads to the device (one per change of the scene)

    RayTracer(Scene &scene);
    ~RayTracer();
    // `tnum` (CPU worker count in the reference) has no meaning on the GPU and
    // is accepted for source compatibility only.
    void start(const uint8_t type, const int8_t tnum = 1);
    void stop();
};
