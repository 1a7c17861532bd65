/*
 * rtg_host.h — C ABI of the host side of librtg: the headless Scene /
 * RayTracer API (raytrace_amd/host) for callers that are not C++, plus the
 * seeded synthetic workloads the benchmarks and parity tests render.
 *
 * The reference exposes these operations only as C++ members of Scene
 * (Scene.h:33-45) and RayTracer (RayTracer.h:51-55); each entry point below
 * names the member it forwards to.  Conventions as in rtg.h.
 */
#ifndef RTG_HOST_H
#define RTG_HOST_H

#include "rtg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtg_host_scene_s *rtg_host_scene_t;

/* Options of the seeded synthetic workloads.  0 / -1 select the workload's
 * own value. */
typedef struct rtg_gen_opts {
    uint64_t seed;     /* 0: the workload's seed (1..5) */
    int32_t width, height;
    int32_t detail;    /* workload 2: icosphere frequency (100 -> 200,000 triangles);
                          workloads 3/4: triangle count; workload 5: terrain divisor */
    int32_t n_spheres; /* -1: default */
    int32_t n_lights;  /* -1: default */
    int32_t index;     /* workloads 3/4: which of the 8 resident scenes */
    int32_t max_level; /* -1: default */
    int32_t pad;
} rtg_gen_opts;

typedef struct rtg_gen_info {
    int32_t width, height, max_level, spp;
    int64_t n_tris;
    int32_t n_objects, n_lights;
} rtg_gen_info;

/* Build workload `config` (1: 640x480 spheres + plane; 2: 1920x1080 displaced
 * 200k-triangle icosphere + 64 spheres; 3: one of the 8 batch scenes; 5: the
 * 4K terrain stress scene) through the Scene API. */
int rtg_host_scene_config(int config, const rtg_gen_opts *opts, rtg_host_scene_t *out, rtg_gen_info *info);
/* One request of the batch and serving workloads (3 and 4). */
typedef struct rtg_gen_request {
    int32_t scene;          /* which of the 8 resident scenes (rtg_gen_opts.index) */
    int32_t width, height;  /* 64x64 .. 1024x768, multiples of 64 */
    int32_t pad;
    double arrival_s;       /* workload 4: Poisson arrival time; workload 3: 0 */
} rtg_gen_request;

/* Request stream of workloads 3 and 4 (seeded splitmix64, u = (x >> 40) 2^-24):
 * request k draws scene floor(8u), width 64 round(16^u), height 64 round(12^u);
 * with rate > 0 arrivals are Poisson at `rate` requests per second. */
int rtg_host_gen_requests(uint64_t seed, int n, double rate, rtg_gen_request *out);

/* Scene::Scene (Scene.cpp:6-67): default camera, EnvLight and material library. */
int rtg_host_scene_new(rtg_host_scene_t *out);
int rtg_host_scene_destroy(rtg_host_scene_t scene);
/* Flattened view of the scene for rtg_scene_create (refreshed by each call;
 * valid until the next call on this handle or its destruction). */
const rtg_scene_desc *rtg_host_scene_desc(rtg_host_scene_t scene);

/* Scene::AddSphere / AddPlane / AddCube / AddBallPlane (Scene.cpp:99-155):
 * returns the object index (>= 0) or < 0. */
int rtg_host_add_sphere(rtg_host_scene_t scene, float radius);
int rtg_host_add_plane(rtg_host_scene_t scene);
int rtg_host_add_cube(rtg_host_scene_t scene, float len);
int rtg_host_add_ballplane(rtg_host_scene_t scene, float radius);
/* Scene::AddModel (Scene.cpp:123-131), OBJ + MTL paths in UTF-8. */
int rtg_host_add_model_obj(rtg_host_scene_t scene, const char *obj_path, const char *mtl_path);
/* Scene::AddLight (Scene.cpp:83-97): returns the light index, 255 when full. */
int rtg_host_add_light(rtg_host_scene_t scene, int type, const float comp[3], const float atte[4]);
/* Scene::MovePos (Scene.cpp:248-280). */
int rtg_host_move_pos(rtg_host_scene_t scene, int type, int num, const float v[3]);
/* Scene::Switch (Scene.cpp:282-301). */
int rtg_host_switch(rtg_host_scene_t scene, int type, int num, int show);
/* Scene::ChgLightComp (Scene.cpp:157-188): type MY_LIGHT_COMPENT (1) weights
 * the ambient / diffuse / specular parts by v[0..2] and renormalises them to
 * the luminance; MY_LIGHT_LUMI (2) scales the luminance by v[3]. */
int rtg_host_chg_light_comp(rtg_host_scene_t scene, int type, int num, const float v[4]);
/* Scene::ChgMtl(num, Normal) (Scene.cpp:202-224): recolour the diffuse term;
 * clr[3] < 0.5 keeps the old r + g + b. */
int rtg_host_recolour(rtg_host_scene_t scene, int num, const float clr[4]);
/* Scene::ChgMtl(num, MtlLiby[lib_index]) (Scene.cpp:190-200). */
int rtg_host_chg_mtl_library(rtg_host_scene_t scene, int num, int lib_index);
/* Scene::Delete (Scene.cpp:226-246): type MY_MODEL_OBJECT (2) or MY_MODEL_LIGHT (1). */
int rtg_host_delete(rtg_host_scene_t scene, int type, int num);
/* Scene::Lights[num] as a flat record. */
int rtg_host_light(rtg_host_scene_t scene, int num, rtg_light *out);
/* Scene::ChgMtl(num, Material) (Scene.cpp:190-200), material given as a flat record. */
int rtg_host_set_material(rtg_host_scene_t scene, int num, const rtg_material *m);
/* Direct writes the reference's callers make to public members:
 * DrawObject::position, Light::position (via SetProperty), Scene::EnvLight,
 * Camera::resize. */
int rtg_host_set_position(rtg_host_scene_t scene, int num, const float pos[3]);
int rtg_host_set_light_position(rtg_host_scene_t scene, int num, const float pos[4]);
int rtg_host_set_env_light(rtg_host_scene_t scene, const float env[4]);
int rtg_host_resize(rtg_host_scene_t scene, int width, int height);

/* RayTracer::start(type, tnum) followed by waiting for isFinish
 * (RayTracer.cpp:614-696), `frames` times on one RayTracer, as the viewer
 * re-renders (main.cpp:323-337): the last frame lands in `out` (2048*2048*3
 * bytes, the reference's fixed buffer); useTime of the last frame in seconds;
 * `uploads` (optional) receives how many times the scene was uploaded. */
int rtg_host_raytrace(rtg_host_scene_t scene, int type, int max_level, int frames, uint8_t *out,
                      double *use_time, uint64_t *uploads);
/* RayTracer::start(type, tnum), then RayTracer::stop() (RayTracer.cpp:698-701)
 * `stop_after_s` seconds later, then waiting for isFinish, as the viewer's
 * close handler does (main.cpp:699-705).  The frame lands in `out` (2048*2048*3
 * bytes): tiles rendered before the stop keep their pixels, the rest stay 127.
 * Returns the frame's status: RTG_E_CANCELLED when the stop reached it.
 * `finished_before_stop` (optional) receives 1 when isFinish was already set
 * at the stop. */
int rtg_host_raytrace_stop(rtg_host_scene_t scene, int type, int max_level, double stop_after_s, uint8_t *out,
                           int *finished_before_stop);

/* One ray against one object of the scene: DrawObject::intersect
 * (3DElement.h:201; Basic3DObject.cpp:135-190, 274-300, 376-410, 486-552;
 * Model.cpp:748-811), after the object's RTPrepare (Model.cpp:402-480), as the
 * reference's renderer calls it for each object in turn.  The ray is
 * Ray(origin, direction, type) with isInside (0 or 0xFF); the hit it must
 * improve on has distance `prev_distance` (1e20: HitRes()) and identity
 * `prev_obj` (what a previous call returned in `out->obj`, or 0).  `min` is
 * the any-hit bound.  `out` receives the returned HitRes; out->obj is the
 * identity the reference compares (the object's, a triangle copy's, or a
 * BallPlane sphere's address), out->hit is 1 when the object replaced the
 * hit, 0 when it returned `prev` unchanged. */
typedef struct rtg_host_ray {
    float origin[4], direction[4];
    int32_t type, inside;
} rtg_host_ray;
typedef struct rtg_host_hit {
    float position[4], normal[4];
    float u, v, distance, rfr;
    int64_t obj;      /* HitRes::obj */
    int64_t object;   /* address of the DrawObject itself, for identities this + k */
    int32_t inside;   /* HitRes::isInside (0 or 0xFF) */
    int32_t hit;
    int32_t has_tex;  /* HitRes::tex != NULL */
    int32_t pad;
} rtg_host_hit;
int rtg_host_intersect(rtg_host_scene_t scene, int num, const rtg_host_ray *ray, int64_t prev_obj,
                       float prev_distance, float min, rtg_host_hit *out);

/* Write a frame as an 8-bit RGB PNG: what the reference's viewer shows of
 * RayTracer::output (main.cpp:207-208, an RGB texture whose row 0 is the
 * bottom).  `stride` bytes per source row (0: 3 * width); bottom_up != 0 flips
 * the rows so that the file's first row is the frame's top.  Returns 0, -1
 * for bad arguments, -2 when the file cannot be written. */
int rtg_host_write_png(const char *path, const uint8_t *rgb, int width, int height, int stride, int bottom_up);

#ifdef __cplusplus
}
#endif
#endif /* RTG_HOST_H */
