// rt_scene.hpp — device-resident scene layout (HBM) and its view.
//
// Layout, per scene:
//   nodes    BVH2, 64 B per node: both children's boxes + links, read as four
//            16-byte loads.  Children are internal node indices (>= 0) or
//            leaves ~(first << 3 | (count - 1)) into `prims`.  The nodes of
//            the top five BFS levels come first (indices < kTopNodes).
//   prims    48 B per primitive in leaf order (3 x float4):
//              triangle: (p0w, gtri) (e1, obj) (e2, 0) — the reference's
//                        clTri values p0 + position, p1 - p0, p2 - p0
//              sphere:   (centre, ~obj) (radius^2, radius, 0, 0) (0)
//   tri_*    per-triangle shading data in scene order, read only for the
//            winning hit: 3 normals, 3 uv, part index, octant-bin mask; and
//            the model-space vertices, read only by the FP64 near-tie re-test.
//   objects  Scene::Objects order; planes, cubes and ballplanes are tested
//            outside the BVH through `brute`.
#pragma once
#include "rt_core.hpp"

namespace rtg {

struct alignas(16) BvhNode {
    float4 xy0;  // child 0: lo.x, hi.x, lo.y, hi.y
    float4 xy1;  // child 1: lo.x, hi.x, lo.y, hi.y
    float4 z01;  // lo.z0, hi.z0, lo.z1, hi.z1
    int4 link;   // child0, child1, -, -
};
static_assert(sizeof(BvhNode) == 64, "64-byte node");

// The nodes of the top five BFS levels are emitted first (indices <
// kTopNodes): traversal kernels that render one scene keep them in LDS.
constexpr int kTopNodes = 31;

RTG_HD bool is_leaf(int c) { return c < 0; }
RTG_HD int leaf_first(int c) { return (~c) >> 3; }
RTG_HD int leaf_count(int c) { return ((~c) & 7) + 1; }
RTG_HD int make_leaf(int first, int count) { return ~((first << 3) | (count - 1)); }

struct DevObject {
    int type, show, mtl, tex;
    V3 pos;
    float radius;
    float rsq;
    int part_begin, part_count;
    int pad0;
    V3 normal, axisx, axisy;  // plane / ballplane frame
    V3 bmin, bmax;            // model: BorderMin/BorderMax (world); cube: min/max (model space)
    int pad1[2];
};

struct DevPart {
    V3 bbmin, bbmax;  // Model::bboxs (world)
    int obj, mtl;
    int first_prim;   // index of the part's first triangle within its model
    int gtri_base;    // global index of the part's first triangle
};

struct DevMaterial {
    V3 ka, kd, ks;
    float shiness, reflect, refract, rfr;
    int tex;
    int pad[2];
};

struct DevLight {
    int type, on;
    V3 pos, la, ld, ls, att;
    float pos_w;  // Light::position.alpha (the material test's light-kind test)
};

struct DevTexture {
    int w, h;
    int off;  // byte offset into the texel pool (B,G,R per texel)
    int pad;
};

struct DevCamera {
    V3 pos, u, v, n;
    float zNear, zFar;  // zFar already multiplied by sqrt(2) (RayTracer.cpp:15)
};

struct SceneView {
    const BvhNode *nodes;
    int n_nodes;  // emitted BVH nodes
    const float4 *prims;
    int root;  // child encoding of the root; 0 with n_prims == 0 means empty
    uint32_t lights_on;  // bit i: light i is switched on (Light::bLight)
    int n_prims;
    const DevObject *objs;
    int n_objs;
    int n_brute;
    const int *brute;
    const DevPart *parts;
    const float4 *tri_nrm;  // 3 per triangle
    const float2 *tri_uv;   // 3 per triangle
    const float4 *tri_pos;  // 3 per triangle: model-space vertices (FP64 near-tie re-test only)
    const int *tri_part;
    const uint8_t *tri_bins;
    const DevMaterial *mtls;
    const DevLight *lights;
    int n_lights;
    int n_texs;
    const DevTexture *texs;
    const uint8_t *texels;
    V3 env;
    DevCamera cam;
};

// Identity of a ray's source (HitRes::obj of the parent hit), compared by the
// self-skip rules: object index, primitive (global triangle / ballplane
// sphere / -1), and for triangles the octant bin of the clTri copy the
// reference recorded (Model.cpp:775,806).
struct SrcId {
    int obj;   // -1: none (primary rays)
    int prim;
    int bin;
};

RTG_HD SrcId no_src() { return SrcId{-1, -1, -1}; }

}  // namespace rtg
