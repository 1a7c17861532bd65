/*
 * rtg.h — C ABI of librtg, the MI355X (gfx950) render path of the RayTrace
 * per-pixel Whitted renderer.
 *
 * The reference has no plugin ABI: its render path sits behind the C++ classes
 * RayTracer (RayTracer.h:16-56), DrawObject (3DElement.h:185-202) and Scene
 * (Scene.h:16-48).  This header is the thin, caller-agnostic boundary those
 * classes are rewired onto (raytrace_amd/host/rt_raytracer.cpp): plain structs,
 * plain pointers and sizes, int status codes.  Every entry point names the
 * reference function it replaces.
 *
 * Conventions
 *   - Return value: RTG_OK (0) on success, < 0 on error; rtg_last_error()
 *     returns a thread-local message for the last failing call on that thread.
 *   - Host buffers are caller-owned; device memory is library-owned unless an
 *     entry point says "device pointer".
 *   - A scene or server handle may be used from one thread at a time.
 *   - Coordinates follow the reference: output row 0 is the bottom image row
 *     (ycur = -H/2), 3 bytes per pixel in R,G,B order, row stride 3*W
 *     (RayTracer.cpp:20-37, 3DElement.cpp:463-468).
 */
#ifndef RTG_H
#define RTG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTG_ABI_VERSION 1

/* status codes */
#define RTG_OK 0
#define RTG_E_INVALID (-1)   /* bad argument */
#define RTG_E_DEVICE (-2)    /* HIP runtime error */
#define RTG_E_NOMEM (-3)     /* device allocation failed / budget exceeded */
#define RTG_E_CANCELLED (-4) /* rtg_cancel() stopped the job */
#define RTG_E_STATE (-5)     /* call not valid in the current state */
#define RTG_E_LIMIT (-6)     /* a documented size limit was exceeded */
#define RTG_E_PENDING 1      /* rtg_poll: job still running */

/* object types: 3DElement.h:4-8 */
#define RTG_OBJ_SPHERE 0x1
#define RTG_OBJ_CUBE 0x2
#define RTG_OBJ_MODEL 0x3
#define RTG_OBJ_PLANE 0x4
#define RTG_OBJ_BALLPLANE 0x5

/* light types: 3DElement.h:13-15 */
#define RTG_LIGHT_PARALLEL 0x1
#define RTG_LIGHT_POINT 0x2
#define RTG_LIGHT_SPOT 0x3

/* render modes: RayTracer.h:5-13 */
#define RTG_MODE_CHECK 0x1
#define RTG_MODE_DEPTHTEST 0x2
#define RTG_MODE_NORMALTEST 0x3
#define RTG_MODE_TEXTURETEST 0x4
#define RTG_MODE_MATERIALTEST 0x5
#define RTG_MODE_SHADOWTEST 0x6
#define RTG_MODE_REFLECTTEST 0x7
#define RTG_MODE_REFRACTTEST 0x8
#define RTG_MODE_RAYTRACE 0x80

/* engines */
#define RTG_ENGINE_AUTO 0
#define RTG_ENGINE_MEGAKERNEL 1 /* one launch, one thread per sample, DFS of the ray tree */
#define RTG_ENGINE_WAVEFRONT 2  /* per-depth extend/shade/occlude kernels over compacted queues */

/* rtg_params.flags */
#define RTG_FLAG_FP64_RETEST 0x1 /* re-test near-tie primary candidates in FP64 */
#define RTG_FLAG_KERNEL_TIMES 0x2 /* time every kernel launch of the render (see rtg_kernel_times) */
#define RTG_FLAG_EAGER 0x4        /* batch: launch kernel by kernel, no captured graph */
#define RTG_FLAG_ONE_STREAM 0x8   /* run every kernel of the frame on the render stream: no side stream for
                                     the shadow rays (callers that must not use a second stream; counter
                                     collection that serialises dispatches) */

/* Hard limits of this build (the reference's own caps are noted beside). */
#define RTG_MAX_LIGHTS 32      /* reference AddLight caps at 8 (Scene.cpp:85-86) */
#define RTG_MAX_LEVEL 15       /* reference maxLevel is a uint8 (RayTracer.h:49) */
#define RTG_MAX_DIM 16384      /* reference output is a fixed 2048x2048 (RayTracer.cpp:603) */

typedef struct rtg_vec4 {
    float x, y, z, w;
} rtg_vec4;

/* Material (3DElement.h:106-120).  `tex` indexes rtg_scene_desc.textures. */
typedef struct rtg_material {
    rtg_vec4 ambient, diffuse, specular, emission;
    float shiness, reflect, refract, rfr;
    int32_t tex; /* -1: no texture (texel colour is white, 3DElement.cpp:432-436) */
    int32_t pad[3];
} rtg_material;

/* Texture (3DElement.h:94-104): w*h texels, 3 bytes each in B,G,R order. */
typedef struct rtg_texture {
    int32_t w, h;
    const uint8_t *bgr;
} rtg_texture;

/* One entry of Scene::Objects (Scene.h:29), flattened. */
typedef struct rtg_object {
    int32_t type;  /* RTG_OBJ_* */
    int32_t show;  /* DrawObject::bShow */
    int32_t mtl;   /* sphere / plane / cube / ballplane: index into materials */
    int32_t tex;   /* plane: index into textures (Plane::tex) */
    rtg_vec4 position;         /* DrawObject::position */
    float radius, radius_sqr;  /* sphere, ballplane (Basic3DObject.h:8,57) */
    int32_t part_begin, part_count; /* model: range of rtg_scene_desc.parts */
    rtg_vec4 normal, axisx, axisy;  /* plane / ballplane frame (Basic3DObject.h:39,43) */
    rtg_vec4 vmin, vmax; /* model: VerMin/VerMax (model space, Model.h:22); cube: min/max */
} rtg_object;

/* One Model part (Model.h:27,30,34): triangles in model space + its border box. */
struct rtg_triangle;
typedef struct rtg_part {
    const struct rtg_triangle *tris; /* the part's triangles (Model::parts[p]) */
    int32_t tri_count; /* <= 32767: clTri::numb is int16 (3DElement.h:125) */
    int32_t mtl;       /* index into materials (part_mtl -> mtls, flattened) */
    rtg_vec4 bmin, bmax; /* Model::borders[2*part], [2*part+1] (model space) */
} rtg_part;

/* Triangle (3DElement.h:128-139): points and normals in model space.  Same
 * 128-byte layout as the reference's Triangle. */
typedef struct rtg_triangle {
    rtg_vec4 p[3];
    rtg_vec4 n[3];
    float uv[6];
    float pad[2];
} rtg_triangle;

/* Light (3DElement.h:204-225). */
typedef struct rtg_light {
    int32_t type; /* RTG_LIGHT_* */
    int32_t on;   /* Light::bLight */
    int32_t pad[2];
    rtg_vec4 position, ambient, diffuse, specular, attenuation;
} rtg_light;

/* Camera (3DElement.h:227-237). */
typedef struct rtg_camera {
    rtg_vec4 position, u, v, n;
    int32_t width, height;
    float fovy, zNear, zFar;
    float pad[3];
} rtg_camera;

/* A whole Scene (Scene.h:16-48), flattened in Scene::Objects order. */
typedef struct rtg_scene_desc {
    const rtg_object *objects;
    int32_t n_objects;
    int32_t n_parts;
    const rtg_part *parts;
    int64_t n_tris; /* sum of parts[i].tri_count */
    const rtg_material *materials;
    int32_t n_materials;
    int32_t n_textures;
    const rtg_texture *textures;
    const rtg_light *lights;
    int32_t n_lights;
    int32_t pad;
    rtg_vec4 env_light; /* Scene::EnvLight */
    rtg_camera camera;  /* Scene::cam */
} rtg_scene_desc;

/* Per-frame render parameters: RayTracer::start(type, tnum) (RayTracer.cpp:614)
 * plus what the reference fixes implicitly. */
typedef struct rtg_params {
    int32_t mode;        /* RTG_MODE_*: RAYTRACE / REFRACTTEST the full tree; REFLECTTEST without refraction;
                            CHECK, DEPTHTEST .. SHADOWTEST the reference's primary-only debug modes */
    int32_t max_level;   /* RayTracer::maxLevel: levels 0..max_level are traced */
    int32_t width, height; /* 0: take the scene camera's */
    int32_t spp;         /* 1: the reference's single pixel-corner sample; 4: 2x2 jittered strata */
    int32_t tile_compat; /* 1: write only whole 64x64 tiles, others stay 127 (RayTracer.cpp:13-21,620) */
    int32_t engine;      /* RTG_ENGINE_* */
    int32_t flags;       /* RTG_FLAG_* */
    uint64_t seed;       /* jitter seed for spp > 1 */
} rtg_params;

/* Ray counts use the reference recursion's semantics (RayTracer.cpp:450-596). */
typedef struct rtg_stats {
    uint64_t rays_primary, rays_shadow, rays_reflect, rays_refract;
    uint64_t rays_total;
    double render_ms; /* device time of the render, HIP events */
    double build_ms;  /* device time of the last BVH build for this scene */
    uint64_t retests_fp64;
    uint64_t graph_replays; /* batch: 1 when the launch sequence was replayed from a captured hipGraph */
} rtg_stats;

typedef struct rtg_scene_s *rtg_scene_t;
typedef struct rtg_job_s *rtg_job_t;

/* ---- library ---------------------------------------------------------- */

/* Replaces RayTracer::RayTracer (RayTracer.cpp:600-606): select the device
 * and create the library's stream.  Idempotent per device. */
int rtg_init(int device);
/* Release every library resource (scenes must be destroyed first). */
int rtg_shutdown(void);
/* Thread-local message of the last failing call (never NULL). */
const char *rtg_last_error(void);
int rtg_abi_version(void);

/* ---- scene residency -------------------------------------------------- */

/* Replaces RayTracer::start's per-frame `RTPrepare` of every shown object
 * (RayTracer.cpp:621-625, Model::RTPrepare Model.cpp:402-480): upload the
 * flattened scene to HBM and build the BVH over spheres and Model triangles on
 * the device.  The scene stays resident until rtg_scene_destroy. */
int rtg_scene_create(const rtg_scene_desc *desc, rtg_scene_t *out);
/* Replaces the effect of Scene mutations (Scene.cpp:157-301) on the render:
 * re-upload `desc` into an existing handle.  The BVH work follows what
 * changed:
 *   - materials, lights, textures, environment light or camera only
 *     (ChgMtl, light edits): the shading tables are re-uploaded, the BVH stays;
 *   - object positions or radii with the same primitive set (MovePos, a model
 *     or sphere moved): the BVH is refitted over its existing topology;
 *   - anything else (Switch, AddX, Delete, other triangles): rebuilt. */
int rtg_scene_update(rtg_scene_t scene, const rtg_scene_desc *desc);
/* rtg_scene_update that always rebuilds the BVH (a fresh tree after many
 * refits have moved objects far from where the tree was built). */
int rtg_scene_rebuild(rtg_scene_t scene, const rtg_scene_desc *desc);
int rtg_scene_destroy(rtg_scene_t scene);
/* Device bytes held by the scene (geometry + BVH + the topology kept for refits). */
int rtg_scene_bytes(rtg_scene_t scene, uint64_t *bytes);
/* Device time (ms) of the scene's last BVH build or refit. */
int rtg_scene_build_ms(rtg_scene_t scene, double *ms);
/* Size of a resident scene and what its updates did. */
typedef struct rtg_scene_info {
    uint64_t bytes;
    double build_ms;
    int32_t n_nodes, n_prims;
    uint64_t rebuilds;        /* updates that rebuilt the BVH */
    uint64_t refits;          /* updates that refitted it */
    uint64_t shading_updates; /* updates that re-uploaded shading tables only */
} rtg_scene_info;
int rtg_scene_stats(rtg_scene_t scene, rtg_scene_info *out);

/* ---- render ----------------------------------------------------------- */

/* Replaces RayTracer::start + parallelRT + RTfrac (RayTracer.cpp:5-45,
 * 450-596, 614-696) for one frame, blocking.  `rgb` receives W*H*3 bytes.
 * `ids` (optional, W*H*2 int32) receives each pixel's primary hit as
 * (object index, triangle index within the model or -1), (-1,-1) for a miss.
 * `tprim` (optional, W*H floats) receives the primary hit distance (1e20 miss). */
int rtg_render(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb,
               int32_t *ids, float *tprim, rtg_stats *stats);

/* Same, with device pointers for the outputs and a caller stream (hipStream_t
 * passed as void*, NULL: the library stream).  Asynchronous with respect to
 * the host; `stats` ray counts are filled when the stream completes only if
 * `stats_dev` is a device pointer to 4 uint64 (primary, shadow, reflect,
 * refract), otherwise not at all. */
int rtg_render_device(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb_dev,
                      int32_t *ids_dev, float *tprim_dev, uint64_t *stats_dev, void *stream);

/* rtg_render_device that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory while they render (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as rtg_render_device with the
 * megakernel engine (engine AUTO or MEGAKERNEL; no FP64 re-test, no kernel
 * timing): every scratch buffer lives in one fixed device array of the
 * library (80.1 MiB per GPU, held until the process ends), set up by rtg_init.
 * Frames up to RTG_MAX_DIM x RTG_MAX_DIM; larger sizes return RTG_E_LIMIT.
 * Calls on one GPU share that array: calls that overlap (two streams, two
 * threads) race.  `stats_dev` NULL: the ray counts stay in the scratch. */
int rtg_render_device_noalloc(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb_dev, int32_t *ids_dev,
                              float *tprim_dev, uint64_t *stats_dev, void *stream);

/* Non-blocking render: RayTracer::start's detached workers + monitor
 * (RayTracer.cpp:664-695).  rtg_poll returns RTG_E_PENDING while running, RTG_OK
 * when finished (isFinish, RayTracer.h:47), RTG_E_CANCELLED after rtg_cancel.
 * Output buffers must stay valid until rtg_job_wait returns. */
int rtg_render_async(rtg_scene_t scene, const rtg_params *params, uint8_t *rgb,
                     int32_t *ids, float *tprim, rtg_job_t *job);
int rtg_poll(rtg_job_t job);
/* Replaces RayTracer::stop (RayTracer.cpp:698-701): cooperative cancel. Tiles
 * not yet rendered keep the value 127, as in the reference. */
int rtg_cancel(rtg_job_t job);
/* Wait for completion and fill `stats` (may be NULL); returns the job's status. */
int rtg_job_wait(rtg_job_t job, rtg_stats *stats);
/* Release a job (waits for it first if it is still running). */
int rtg_job_free(rtg_job_t job);

/* ---- kernel timing ---------------------------------------------------- */

/* Per-kernel device time of the renders issued with RTG_FLAG_KERNEL_TIMES
 * since the previous call: HIP events recorded on the render's stream around
 * every launch.  Waits for those events.  `items` counts the work units the
 * launches processed (rays, shadow rays, tree nodes or pixels).  No
 * counterpart in the reference, which has no device. */
typedef struct rtg_kernel_time {
    char name[24];
    uint32_t launches;
    uint32_t pad;
    uint64_t items;
    double ms;
} rtg_kernel_time;
int rtg_kernel_times(rtg_kernel_time *out, int max, int *n);

/* ---- batched rendering (many requests, one persistent launch) --------- */

typedef struct rtg_request {
    rtg_scene_t scene;
    rtg_params params;
    uint8_t *rgb;   /* host W*H*3 */
    int32_t *ids;   /* optional host W*H*2 */
} rtg_request;

/* Render n requests (n <= 4096) of uneven size and scene together, as one
 * wavefront: all their rays share one global queue per depth, processed by
 * persistent kernels; the launch sequence is captured into a hipGraph and
 * replayed for later batches of the same padded shape (unless a request sets
 * RTG_FLAG_EAGER).  Each request's result is identical to rendering it alone
 * with rtg_render.  `stats` sums the ray counts of all requests.  No
 * counterpart in the reference, which renders one frame per start(). */
int rtg_render_batch(const rtg_request *reqs, int n, rtg_stats *stats);

/* ---- launch tuning of the wavefront engine ----------------------------- */

/* Tile and queue sizes of the wavefront engine (frames and batches).  None of
 * them changes any pixel: only the order in which rays are traced and how the
 * persistent kernels split the queues.  No counterpart in the reference,
 * whose CPU workers take fixed 64x64 tiles (RayTracer.cpp:13-41). */
typedef struct rtg_tuning {
    int32_t tile_w;   /* width of the pixel block whose 64 primary rays one wave traces:
                         4, 8, 16 or 32 (height 64 / tile_w); default 8 */
    int32_t grab;     /* 64-item queue chunks a wave takes per atomic: 1, 2 or 4; default 1 */
    int32_t grid_pct; /* persistent grid of the traversal kernels, percent of the blocks
                         the GPU holds at once: 25..100; default 100 */
    int32_t reserved; /* 0 */
} rtg_tuning;

/* One timed candidate of rtg_autotune. */
typedef struct rtg_tune_trial {
    rtg_tuning tuning;
    double device_ms; /* median device time of the requests under this tuning */
} rtg_tune_trial;

/* The tuning in use (rtg_init installs the defaults). */
int rtg_get_tuning(rtg_tuning *out);
/* Install a tuning; RTG_E_INVALID for a value outside the ranges above. */
int rtg_set_tuning(const rtg_tuning *t);
/* Autotuner: times the wavefront engine on `reqs` (one request: the frame
 * path; several: one batch, as rtg_render_batch renders it) under candidate
 * tunings, by coordinate descent over tile_w, grab and grid_pct from the
 * tuning in use, each candidate the median of `reps` timed runs after one
 * warmup run.  Installs the fastest and writes it to `best`; the timed
 * candidates go to `trials` (up to max_trials; *n_trials = how many were
 * timed).  Outputs stay on the device. */
int rtg_autotune(const rtg_request *reqs, int n, int reps, rtg_tuning *best, rtg_tune_trial *trials,
                 int max_trials, int *n_trials);

/* ---- serving loop ----------------------------------------------------- */

/* One request of a serving run: a frame of a resident scene, arriving at
 * `arrival_s` seconds after the run starts; RGB lands in `rgb` (host,
 * W*H*3).  No counterpart in the reference, which renders one frame per
 * start() call and has no request loop. */
typedef struct rtg_serve_request {
    rtg_scene_t scene;
    rtg_params params;
    uint8_t *rgb;
    double arrival_s;
    double done_s;   /* out: seconds after the start when the RGB was in host memory */
    int32_t status;  /* out: RTG_OK or an error code */
    int32_t batch;   /* out: which batch rendered it */
} rtg_serve_request;

#define RTG_SERVE_BATCHED 0 /* continuous batching over concurrent slots */
#define RTG_SERVE_PLAIN 1   /* one render per request, first come first served */

typedef struct rtg_serve_opts {
    int32_t mode;        /* RTG_SERVE_* */
    int32_t slots;       /* batches in flight at once, each on its own stream (0: 2) */
    int32_t max_batch;   /* requests per batch (0: 256, at most 4096) */
    int32_t graph_after; /* capture a batch shape (of 8 or more requests) into a hipGraph at its n-th
                            launch (0: 3; < 0: never) */
    uint64_t hbm_budget; /* queue bytes one batch may reserve (0: 4 GiB); admission stops there */
} rtg_serve_opts;

typedef struct rtg_serve_stats {
    double wall_s;                       /* first arrival to last completion */
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency: arrival -> RGB in host memory */
    double mrays_per_s, req_per_s;
    uint64_t rays_total;                 /* reference-semantic rays of all requests */
    int32_t batches, max_in_flight;      /* batches launched; most requests on the GPU at once */
} rtg_serve_stats;

/* Serve n requests (open loop: each is admitted once its arrival time has
 * passed).  RTG_SERVE_BATCHED: whenever a slot is free, every request that
 * has arrived and is waiting is admitted into one batch, up to max_batch and
 * the HBM budget, and rendered as one wavefront (rtg_render_batch's engine)
 * while the other slots' batches run; requests that arrive meanwhile wait
 * for the next free slot.  RTG_SERVE_PLAIN: the same arrivals rendered one
 * request at a time, each as its own frame.  Each request's RGB is identical
 * to rtg_render of it alone. */
int rtg_serve(rtg_serve_request *reqs, int n, const rtg_serve_opts *opts, rtg_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* RTG_H */
