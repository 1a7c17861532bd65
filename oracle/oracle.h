/*
 * oracle.h — CPU oracle of the RayTrace per-pixel render path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (raytrace_amd/, librtg)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * Modes
 *   ORACLE_REF32  The reference algorithm restated in FP32 with its operation
 *                 order: per-object scan (RayTracer.cpp:458-465), Model octant
 *                 binning of world-space clTri copies (Model.cpp:402-480),
 *                 BorderTest / BorderTestEx culling (Basic3DObject.cpp:44-81,
 *                 Model.cpp:482-664), Moller-Trumbore (Model.cpp:714-744),
 *                 recursive Whitted shading (RayTracer.cpp:450-596).  This is
 *                 also the "port" CPU baseline that is timed.
 *   ORACLE_FP64   Same recursion and formulas in double precision, intersection
 *                 through the canonical BVH (no octant culling, self-skip by
 *                 primitive identity).  Used for pixel-error statistics.
 *
 * Parity pins: see tests/test_oracle_pins.py.  The reference ships no golden
 * vectors, tests or fixtures, and cannot be built here (Win32 / GL / MSVC).
 */
#ifndef RTG_ORACLE_H
#define RTG_ORACLE_H

#include "../include/rtg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_REF32 0
#define ORACLE_FP64 1

/* per-pixel flag bits */
#define ORACLE_FLAG_NEAR_TIE 0x1  /* primary hit inside the near-tie band */
#define ORACLE_FLAG_CUT_EDGE 0x2  /* primary distance within 1e-5 (relative) of zNear / zFar */

typedef struct oracle_stats {
    uint64_t rays_primary, rays_shadow, rays_reflect, rays_refract;
    uint64_t pixels;
    double seconds; /* wall time of the render */
    int32_t threads;
    int32_t pad;
} oracle_stats;

/* Canonical-BVH traffic counts of a frame's ray set (binned SAH with 16 bins,
 * BVH2, <= 4 primitives per leaf, FP64 build). */
typedef struct oracle_traffic {
    /* every traced ray: closest-hit rays + shadow rays that can change the
     * colour (a shadow ray toward a light with N.l <= 0 and N.h <= 0 is not
     * traced by the device; its outcome cannot change the pixel) */
    uint64_t rays;
    uint64_t node_visits;  /* internal-node box-pair tests */
    uint64_t tri_tests;
    uint64_t sph_tests;
    uint64_t n_nodes, n_prims;
    /* the shadow (any-hit) part of the above */
    uint64_t shadow_rays, shadow_node_visits, shadow_tri_tests, shadow_sph_tests;
    uint64_t shadow_skipped; /* reference shadow rays whose outcome cannot matter */
} oracle_traffic;

typedef struct oracle_scene_s *oracle_scene_t;

/* Model::RTPrepare for every shown model + canonical BVH build. */
int oracle_scene_create(const rtg_scene_desc *desc, oracle_scene_t *out);
void oracle_scene_destroy(oracle_scene_t s);

/* Render a frame (or the listed 64x64 tiles of the full-frame tiling,
 * tile = by * ceil(W/64) + bx).  Pixels not rendered are left untouched.
 * Outputs are optional except rgb: rgb W*H*3; ids W*H*2 (object, triangle
 * index within the model / sub-sphere index, -1 none); tprim W*H; rgbf
 * W*H*3 floats (colour before quantisation); flags W*H bytes. */
int oracle_render(oracle_scene_t s, const rtg_params *params, int mode, int threads, const int32_t *tiles,
                  int n_tiles, uint8_t *rgb, int32_t *ids, float *tprim, float *rgbf, uint8_t *flags,
                  oracle_stats *stats);

/* Trace the frame's full ray set (REF32 semantics) through the canonical BVH
 * and count node visits and primitive tests (closest-hit traversal, nearer
 * child first; shadow rays any-hit). */
int oracle_traffic_count(oracle_scene_t s, const rtg_params *params, int threads, const int32_t *tiles,
                         int n_tiles, oracle_traffic *out);

/* Unit entry points used by the pinning tests (REF32 arithmetic).  Vectors
 * are float[4]. */
float oracle_dot(const float a[4], const float b[4]);
void oracle_cross(const float a[4], const float b[4], float out[4]);
void oracle_normalize(const float a[4], float out[4]);
float oracle_tri_test(const float o[4], const float d[4], const float e1[4], const float e2[4], const float p0[4],
                      float bary[3]);
float oracle_sphere_test(const float o[4], const float d[4], const float c[4], float radius, float *t_inside);
float oracle_plane_test(const float o[4], const float d[4], const float p0[4], const float n[4]);
float oracle_border_test(const float o[4], const float d[4], const float bmin[4], const float bmax[4], float *tmax);
float oracle_border_test_ex(const float o[4], const float d[4], const float bmin[4], const float bmax[4],
                            int32_t mask[8]);
void oracle_raygen(const rtg_camera *cam, int width, int height, int px, int py, float o[4], float d[4]);
uint8_t oracle_quantize(float c);
float oracle_plane_hit(const float o[4], const float d[4], const float p0[4], const float n[4], const float ax[4],
                       const float ay[4], float uv[2]);
float oracle_box_hit(const float o[4], const float d[4], const float pos[4], const float bmin[4], const float bmax[4],
                     float normal[4]);
float oracle_ballplane_hit(const float o[4], const float d[4], const float pos[4], const float ax[4],
                           const float ay[4], float radius, float mtl_rfr, int src, int inside, int type,
                           int32_t out_i[2], float out_f[4]);
void oracle_texel(const rtg_texture *tex, float u, float v, float rgb[3]);

#ifdef __cplusplus
}
#endif
#endif
