// rt_scene.hpp — device-resident scene layout (HBM) and its view.
//
// Layout, per scene:
//   nodes    4-wide BVH, 64 B per node (Bvh4Node), read as four 16-byte
//            loads: up to four children, their boxes quantised to 8 bits per
//            bound on a power-of-two grid anchored at the node's own box, and
//            their links.  Children are internal node indices (>= 0) or leaves
//            ~(first << 3 | (count - 1)) into `prims`.  The top levels come
//            first (indices < kTopNodes).
//   prims    48 B per primitive in leaf order (3 x float4):
//              triangle: (p0w, gtri) (e1, obj) (e2, 0) — the reference's
//                        clTri values p0 + position, p1 - p0, p2 - p0
//              sphere:   (centre, ~obj) (radius^2, radius, 0, 0) (0)
//   tri_*    per-triangle shading data in scene order, read only for the
//            winning hit: 3 normals, 3 uv, part index, material index,
//            octant-bin mask; and the model-space vertices, read only by the
//            FP64 near-tie re-test.
//   objects  Scene::Objects order; planes, cubes and ballplanes are tested
//            outside the BVH through `brute`.
#pragma once
#include "rt_core.hpp"

namespace rtg {

// Binary node of the build (both children's boxes and links); the renderer
// walks the 4-wide nodes made from it.
struct alignas(16) BvhNode {
    float4 xy0;  // child 0: lo.x, hi.x, lo.y, hi.y
    float4 xy1;  // child 1: lo.x, hi.x, lo.y, hi.y
    float4 z01;  // lo.z0, hi.z0, lo.z1, hi.z1
    int4 link;   // child0, child1, -, -
};
static_assert(sizeof(BvhNode) == 64, "64-byte node");

// 4-wide node.  Child k's box on axis a is [org_a + qlo * s_a, org_a + qhi * s_a]
// (one float product, exact, and one rounded sum), s_a = 2^(e_a - 127); the
// bytes are chosen so that the decoded box contains the child's exact box.
struct alignas(16) Bvh4Node {
    float4 org;  // x, y, z: origin (the node box's lower corner); w: bits e_x | e_y << 8 | e_z << 16 | count << 24
    uint4 qxy;   // byte k of x: qlo.x of child k; y: qhi.x; z: qlo.y; w: qhi.y
    uint4 qz;    // byte k of x: qlo.z; y: qhi.z; z, w: 0
    int4 link;   // child k's link (slots >= count: 0)
};
static_assert(sizeof(Bvh4Node) == 64, "64-byte node");

// Internal levels of a 4-wide tree the traversal stack holds: a walk leaves at
// most 3 entries per level.  A Karras tree over 62-bit keys has under 63
// internal binary levels, so its even-depth widening stays within this.
constexpr int kMaxLevels4 = 32;

// The first nodes are the top levels of the tree (indices < kTopNodes):
// traversal kernels that render one scene keep them in LDS.
constexpr int kTopNodes = 31;

RTG_HD bool is_leaf(int c) { return c < 0; }
RTG_HD int leaf_first(int c) { return (~c) >> 3; }
RTG_HD int leaf_count(int c) { return ((~c) & 7) + 1; }
RTG_HD int make_leaf(int first, int count) { return ~((first << 3) | (count - 1)); }

RTG_HD int node4_count(const Bvh4Node &n) { return (int)(__builtin_bit_cast(uint32_t, n.org.w) >> 24); }
RTG_HD float pow2f(int e) { return __builtin_bit_cast(float, (uint32_t)e << 23); }  // 2^(e - 127), e in 1..254
// org + q s.  q s is exact (8 bits times a power of two), so a fused
// multiply-add gives the same value as the product and the sum, one rounding.
RTG_HD float dequant(float org, uint32_t q, float s) { return fmaf((float)q, s, org); }
// Child k's decoded box: lo/hi per axis.
RTG_HD void node4_box(const Bvh4Node &n, int k, float &lx, float &hx, float &ly, float &hy, float &lz, float &hz) {
    const uint32_t e = __builtin_bit_cast(uint32_t, n.org.w);
    const float sx = pow2f(e & 0xFF), sy = pow2f((e >> 8) & 0xFF), sz = pow2f((e >> 16) & 0xFF);
    const int sh = 8 * k;
    lx = dequant(n.org.x, (n.qxy.x >> sh) & 0xFF, sx), hx = dequant(n.org.x, (n.qxy.y >> sh) & 0xFF, sx);
    ly = dequant(n.org.y, (n.qxy.z >> sh) & 0xFF, sy), hy = dequant(n.org.y, (n.qxy.w >> sh) & 0xFF, sy);
    lz = dequant(n.org.z, (n.qz.x >> sh) & 0xFF, sz), hz = dequant(n.org.z, (n.qz.y >> sh) & 0xFF, sz);
}

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
    const Bvh4Node *nodes;
    int n_nodes;  // 4-wide nodes
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
    const int *tri_mtl;  // material of the triangle's part (read beside tri_part)
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
