// rtg_bvh.hip — scene upload and device LBVH build (steps in rt_bvh.hpp).
#include <chrono>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rt_bvh.hpp"
#include "rt_prep.hpp"
#include "rtg_internal.hpp"

namespace rtg {

namespace {

constexpr int kBlock = 256;

__global__ void k_setup_tris(const rtg_triangle *raw, int n, const int *tri_part, const DevPart *parts,
                             const float4 *part_mid, const DevObject *objs, float4 *prims, Box6 *boxes,
                             float4 *nrm, float2 *uv, float4 *pos, uint8_t *bins) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int pi = tri_part[g];
    const DevPart &P = parts[pi];
    const float4 m = part_mid[pi];
    prim_setup_tri(raw[g], objs[P.obj].pos, V3{m.x, m.y, m.z}, g, P.obj, prims + 3 * (size_t)g, nrm + 3 * (size_t)g,
                   uv + 3 * (size_t)g, pos + 3 * (size_t)g, bins[g], boxes[g]);
}

__global__ void k_setup_spheres(const int *sph, int n, int base, const DevObject *objs, float4 *prims, Box6 *boxes) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DevObject &o = objs[sph[k]];
    prim_setup_sphere(o.pos, o.radius, o.rsq, sph[k], prims + 3 * (size_t)(base + k), boxes[base + k]);
}

// Centroid bounds: per-block partials, then one block folds them.
__global__ void k_bounds_partial(const Box6 *boxes, int n, float *partial) {
    __shared__ float s[6][kBlock];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
            mn[k] = fminf(mn[k], c);
            mx[k] = fmaxf(mx[k], c);
        }
    for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = mn[k], s[3 + k][threadIdx.x] = mx[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w /= 2) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_bounds_final(const float *partial, int nb, float *cb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) {
            r[k] = fminf(r[k], partial[6 * b + k]);
            r[3 + k] = fmaxf(r[3 + k], partial[6 * b + 3 + k]);
        }
    for (int k = 0; k < 6; ++k) cb[k] = r[k];
}

__global__ void k_morton(const Box6 *boxes, int n, const float *cb, uint64_t *keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float cmin[3], cinv[3];
    morton_grid(cb, cmin, cinv);
    keys[i] = morton_key(boxes[i], cmin, cinv, i);
}

__global__ void k_karras(const uint64_t *keys, int n, int *left, int *right, int *first, int *last, int *parent_int,
                         int *parent_leaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int l, r, f, la;
    karras_node(keys, n, i, l, r, f, la);
    left[i] = l, right[i] = r, first[i] = f, last[i] = la;
    if (l < 0)
        parent_leaf[~l] = i;
    else
        parent_int[l] = i;
    if (r < 0)
        parent_leaf[~r] = i;
    else
        parent_int[r] = i;
}

__device__ Box6 load_box_coherent(const Box6 *b) {
    Box6 r;
    const int *p = reinterpret_cast<const int *>(b);
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = as_float(__hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        r.hi[k] = as_float(__hip_atomic_load(p + 3 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    return r;
}
__device__ void store_box_coherent(Box6 *b, const Box6 &v) {
    int *p = reinterpret_cast<int *>(b);
    for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(p + k, as_int(v.lo[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 3 + k, as_int(v.hi[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Bottom-up box union: the second child to arrive at a node computes it.
__global__ void k_refit(const uint64_t *keys, int n, const Box6 *leaf_boxes, const int *left, const int *right,
                        const int *parent_int, const int *parent_leaf, int *flags, Box6 *ibox) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = parent_leaf[k];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&flags[p], 1) == 0) return;
        __threadfence();
        const int l = left[p], r = right[p];
        const Box6 bl = l < 0 ? leaf_boxes[(uint32_t)keys[~l]] : load_box_coherent(&ibox[l]);
        const Box6 br = r < 0 ? leaf_boxes[(uint32_t)keys[~r]] : load_box_coherent(&ibox[r]);
        store_box_coherent(&ibox[p], box_union(bl, br));
        p = p == 0 ? -1 : parent_int[p];
    }
}

__global__ void k_emit_flags(const int *first, const int *last, int n_int, int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_int) return;
    flag[i] = last[i] - first[i] + 1 > kLeafMax ? 1 : 0;
}

__global__ void k_emit(const uint64_t *keys, int n_int, const Box6 *leaf_boxes, const int *left, const int *right,
                       const int *first, const int *last, const Box6 *ibox, const int *newidx, BvhNode *nodes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_int) return;
    if (last[i] - first[i] + 1 <= kLeafMax) return;
    const int l = left[i], r = right[i];
    const Box6 bl = l < 0 ? leaf_boxes[(uint32_t)keys[~l]] : ibox[l];
    const Box6 br = r < 0 ? leaf_boxes[(uint32_t)keys[~r]] : ibox[r];
    nodes[newidx[i]] = make_node(bl, br, emit_link(l, first, last, newidx), emit_link(r, first, last, newidx));
}

__global__ void k_reorder(const uint64_t *keys, int n, const float4 *src, float4 *dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t id = (uint32_t)keys[k];
    for (int j = 0; j < 3; ++j) dst[3 * (size_t)k + j] = src[3 * (size_t)id + j];
}

// Top BFS levels first (top_first, rt_bvh.hpp); one thread.
__global__ void k_top_first(const int *left, const int *right, const int *first, const int *last, int n_int,
                            int *newidx) {
    if (blockIdx.x == 0 && threadIdx.x == 0) top_first(left, right, first, last, n_int, newidx);
}

// Widening into 4-wide nodes (rt_bvh.hpp): parents of the binary nodes,
// their depth parity, then one 4-wide node per binary node at even depth.
__global__ void k_parent2(const BvhNode *nodes2, int n2, int *parent2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const int4 l = nodes2[i].link;
    if (l.x >= 0) parent2[l.x] = i;
    if (l.y >= 0) parent2[l.y] = i;
}

__global__ void k_even_flags(const int *parent2, int n2, int *flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) flags[i] = node2_even(parent2, i) ? 1 : 0;
}

__global__ void k_widen(const BvhNode *nodes2, const int *flags, const int *idx4, int n2, Bvh4Node *nodes4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2 || !flags[i]) return;
    nodes4[idx4[i]] = widen_node(nodes2, idx4, i);
}

__global__ void k_fill_i32(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

V3 v3of(const rtg_vec4 &a) { return V3{a.x, a.y, a.z}; }

template <class T>
int upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    if (ensure(b, n) != RTG_OK) return RTG_E_NOMEM;
    if (!v.empty()) RTG_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return RTG_OK;
}

// The 4-wide tree of the n2 binary nodes in nodes2 (root 0) into S.nodes;
// returns its node count in *n4.
int widen_tree(const DevBuf &nodes2, int n2, SceneRes &S, hipStream_t stream, int *n4) {
    DevBuf parent2, flags, idx4, scan_tmp;
    if (ensure(parent2, n2 * sizeof(int)) || ensure(flags, n2 * sizeof(int)) || ensure(idx4, n2 * sizeof(int)))
        return fail(RTG_E_NOMEM, "device allocation failed (4-wide nodes)");
    k_fill_i32<<<nblk(n2), kBlock, 0, stream>>>(parent2.as<int>(), n2, -1);
    k_parent2<<<nblk(n2), kBlock, 0, stream>>>(nodes2.as<BvhNode>(), n2, parent2.as<int>());
    k_even_flags<<<nblk(n2), kBlock, 0, stream>>>(parent2.as<int>(), n2, flags.as<int>());
    size_t tmp_bytes = 0;
    RTG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, flags.as<int>(), idx4.as<int>(), 0, (size_t)n2,
                                    rocprim::plus<int>(), stream));
    if (ensure(scan_tmp, std::max<size_t>(tmp_bytes, 16))) return fail(RTG_E_NOMEM, "scan temp");
    RTG_HIP(rocprim::exclusive_scan(scan_tmp.p, tmp_bytes, flags.as<int>(), idx4.as<int>(), 0, (size_t)n2,
                                    rocprim::plus<int>(), stream));
    int last_idx = 0, last_flag = 0;
    RTG_HIP(hipMemcpyAsync(&last_idx, idx4.as<int>() + (n2 - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
    RTG_HIP(hipMemcpyAsync(&last_flag, flags.as<int>() + (n2 - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
    RTG_HIP(hipStreamSynchronize(stream));
    *n4 = last_idx + last_flag;
    if (ensure(S.nodes, std::max(*n4, 1) * sizeof(Bvh4Node))) return fail(RTG_E_NOMEM, "bvh nodes");
    k_widen<<<nblk(n2), kBlock, 0, stream>>>(nodes2.as<BvhNode>(), flags.as<int>(), idx4.as<int>(), n2,
                                             S.nodes.as<Bvh4Node>());
    RTG_HIP(hipGetLastError());
    RTG_HIP(hipStreamSynchronize(stream));  // the temporaries are freed on return
    return RTG_OK;
}

}  // namespace

// Device bytes a resident scene holds: render data plus the kept topology.
static size_t scene_bytes(const SceneRes &S) {
    return S.nodes.bytes + S.prims.bytes + S.objs.bytes + S.brute.bytes + S.parts.bytes + S.tri_nrm.bytes +
           S.tri_uv.bytes + S.tri_pos.bytes + S.tri_part.bytes + S.tri_mtl.bytes + S.tri_bins.bytes + S.mtls.bytes + S.lights.bytes +
           S.texs.bytes + S.texels.bytes + S.t_keys.bytes + S.t_left.bytes + S.t_right.bytes + S.t_first.bytes +
           S.t_last.bytes + S.t_pint.bytes + S.t_pleaf.bytes + S.t_newidx.bytes;
}

// Lights switched on and the most child rays one hit can spawn (reflection
// and/or refraction) over the materials of shown objects: the latter bounds
// the wavefront queues per level.
static void shading_state(SceneRes &S, const rtg_scene_desc &d) {
    S.lights_on = 0;
    for (int i = 0; i < d.n_lights; ++i)
        if (d.lights[i].on) S.lights_on |= 1u << i;
    auto fan_of = [&](int m) {
        if (m < 0 || m >= d.n_materials) return 0;
        return (d.materials[m].reflect > 0.01f ? 1 : 0) + (d.materials[m].refract > 0.01f ? 1 : 0);
    };
    S.max_fanout = 0;
    for (int i = 0; i < d.n_objects; ++i) {
        const rtg_object &o = d.objects[i];
        if (!o.show) continue;
        if (o.type == RTG_OBJ_MODEL) {
            for (int k = 0; k < o.part_count; ++k) S.max_fanout = std::max(S.max_fanout, fan_of(d.parts[o.part_begin + k].mtl));
        } else {
            S.max_fanout = std::max(S.max_fanout, fan_of(o.mtl));
        }
    }

}

int scene_build(SceneRes &S, const rtg_scene_desc &d, hipStream_t stream, SceneRes *donor) {
    ScenePrep prep;
    const std::string err = prepare_scene(d, prep);
    if (!err.empty()) return fail(RTG_E_INVALID, err);
    const std::vector<DevObject> &objs = prep.objs;
    const std::vector<int> &brute = prep.brute, &spheres = prep.spheres, &tri_part = prep.tri_part;
    const std::vector<DevPart> &parts = prep.parts;
    const std::vector<float4> &part_mid = prep.part_mid;
    const std::vector<DevMaterial> &mtls = prep.mtls;
    const std::vector<DevLight> &lights = prep.lights;
    const std::vector<DevTexture> &texs = prep.texs;
    const std::vector<uint8_t> &texels = prep.texels;
    const int64_t n_tris = prep.n_tris;

    // ---- upload
    int st;
    if ((st = upload(S.objs, objs, stream)) || (st = upload(S.brute, brute, stream)) ||
        (st = upload(S.parts, parts, stream)) || (st = upload(S.mtls, mtls, stream)) ||
        (st = upload(S.lights, lights, stream)) || (st = upload(S.texs, texs, stream)) ||
        (st = upload(S.texels, texels, stream)) || (st = upload(S.tri_part, tri_part, stream)) ||
        (st = upload(S.tri_mtl, prep.tri_mtl, stream)))
        return st;
    const int n = (int)(n_tris + (int64_t)spheres.size());
    // refit: the donor's tree has the same primitive count (and the caller
    // checked the same primitive order); its topology moves into S
    const bool refit = donor != nullptr && donor->topo_n == n && n > kLeafMax;
    if (refit) {
        DevBuf *mine[] = {&S.t_keys, &S.t_left, &S.t_right, &S.t_first, &S.t_last, &S.t_pint, &S.t_pleaf, &S.t_newidx};
        DevBuf *theirs[] = {&donor->t_keys, &donor->t_left,  &donor->t_right, &donor->t_first,
                            &donor->t_last, &donor->t_pint, &donor->t_pleaf, &donor->t_newidx};
        for (int k = 0; k < 8; ++k) std::swap(mine[k]->p, theirs[k]->p), std::swap(mine[k]->bytes, theirs[k]->bytes);
        S.topo_n = donor->topo_n, S.topo_emit = donor->topo_emit;
        donor->topo_n = 0, donor->topo_emit = 0;
    }
    DevBuf raw, mid, sph, prims_u, boxes, keys, flags, ibox, partial, cb, sort_tmp, nodes2;
    DevBuf &keys2 = S.t_keys, &left = S.t_left, &right = S.t_right, &first = S.t_first, &last = S.t_last,
           &pint = S.t_pint, &pleaf = S.t_pleaf, &newidx = S.t_newidx;
    if ((st = upload(mid, part_mid, stream)) || (st = upload(sph, spheres, stream))) return st;
    if (ensure(raw, std::max<size_t>(n_tris, 1) * sizeof(rtg_triangle))) return RTG_E_NOMEM;
    {
        size_t off = 0;
        for (const int pi : prep.part_order) {
            const rtg_triangle *src = prep.part_tris[pi];
            const int cnt = d.parts[pi].tri_count;
            if (src == nullptr || cnt == 0) continue;
            RTG_HIP(hipMemcpyAsync(raw.as<char>() + off, src, sizeof(rtg_triangle) * cnt, hipMemcpyHostToDevice, stream));
            off += sizeof(rtg_triangle) * cnt;
        }
    }
    const size_t nt = std::max<int64_t>(n_tris, 1);
    if (ensure(S.tri_nrm, nt * 3 * sizeof(float4)) || ensure(S.tri_uv, nt * 3 * sizeof(float2)) ||
        ensure(S.tri_pos, nt * 3 * sizeof(float4)) ||
        ensure(S.tri_bins, nt) || ensure(prims_u, std::max(n, 1) * 3 * sizeof(float4)) ||
        ensure(boxes, std::max(n, 1) * sizeof(Box6)))
        return fail(RTG_E_NOMEM, "device allocation failed (scene)");

    hipEvent_t e0, e1;
    RTG_HIP(hipEventCreate(&e0));
    RTG_HIP(hipEventCreate(&e1));
    RTG_HIP(hipEventRecord(e0, stream));
    if (n_tris > 0)
        k_setup_tris<<<nblk(n_tris), kBlock, 0, stream>>>(raw.as<rtg_triangle>(), (int)n_tris, S.tri_part.as<int>(),
                                                          S.parts.as<DevPart>(), mid.as<float4>(), S.objs.as<DevObject>(),
                                                          prims_u.as<float4>(), boxes.as<Box6>(), S.tri_nrm.as<float4>(),
                                                          S.tri_uv.as<float2>(), S.tri_pos.as<float4>(),
                                                          S.tri_bins.as<uint8_t>());
    if (!spheres.empty())
        k_setup_spheres<<<nblk(spheres.size()), kBlock, 0, stream>>>(sph.as<int>(), (int)spheres.size(), (int)n_tris,
                                                                     S.objs.as<DevObject>(), prims_u.as<float4>(),
                                                                     boxes.as<Box6>());
    int root = 0, n_emit = 0;
    if (refit) {
        // same topology: leaf boxes from the new primitives, bottom-up union,
        // nodes re-emitted with the same links, primitives into leaf order
        if (ensure(flags, n * sizeof(int)) || ensure(ibox, n * sizeof(Box6)))
            return fail(RTG_E_NOMEM, "device allocation failed (bvh refit)");
        const uint64_t *sk = keys2.as<uint64_t>();
        k_fill_i32<<<nblk(n), kBlock, 0, stream>>>(flags.as<int>(), n, 0);
        k_refit<<<nblk(n), kBlock, 0, stream>>>(sk, n, boxes.as<Box6>(), left.as<int>(), right.as<int>(), pint.as<int>(),
                                                pleaf.as<int>(), flags.as<int>(), ibox.as<Box6>());
        n_emit = S.topo_emit;
        if (ensure(nodes2, std::max(n_emit, 1) * sizeof(BvhNode))) return fail(RTG_E_NOMEM, "bvh nodes");
        k_emit<<<nblk(n - 1), kBlock, 0, stream>>>(sk, n - 1, boxes.as<Box6>(), left.as<int>(), right.as<int>(),
                                                   first.as<int>(), last.as<int>(), ibox.as<Box6>(), newidx.as<int>(),
                                                   nodes2.as<BvhNode>());
        if (ensure(S.prims, (size_t)n * 3 * sizeof(float4))) return fail(RTG_E_NOMEM, "prims");
        k_reorder<<<nblk(n), kBlock, 0, stream>>>(sk, n, prims_u.as<float4>(), S.prims.as<float4>());
        root = 0;
    } else if (n > kLeafMax) {
        const int nb = 512;
        if (ensure(partial, nb * 6 * sizeof(float)) || ensure(cb, 6 * sizeof(float)) ||
            ensure(keys, n * sizeof(uint64_t)) || ensure(keys2, n * sizeof(uint64_t)) ||
            ensure(left, n * sizeof(int)) || ensure(right, n * sizeof(int)) || ensure(first, n * sizeof(int)) ||
            ensure(last, n * sizeof(int)) || ensure(pint, n * sizeof(int)) || ensure(pleaf, n * sizeof(int)) ||
            ensure(flags, n * sizeof(int)) || ensure(ibox, n * sizeof(Box6)) || ensure(newidx, n * sizeof(int)))
            return fail(RTG_E_NOMEM, "device allocation failed (bvh build)");
        k_bounds_partial<<<nb, kBlock, 0, stream>>>(boxes.as<Box6>(), n, partial.as<float>());
        k_bounds_final<<<1, 64, 0, stream>>>(partial.as<float>(), nb, cb.as<float>());
        k_morton<<<nblk(n), kBlock, 0, stream>>>(boxes.as<Box6>(), n, cb.as<float>(), keys.as<uint64_t>());
        size_t tmp_bytes = 0;
        RTG_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (size_t)n, 0, 62,
                                         stream));
        if (ensure(sort_tmp, std::max<size_t>(tmp_bytes, 16))) return fail(RTG_E_NOMEM, "sort temp");
        RTG_HIP(rocprim::radix_sort_keys(sort_tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (size_t)n, 0,
                                         62, stream));
        const uint64_t *sk = keys2.as<uint64_t>();
        k_fill_i32<<<nblk(n), kBlock, 0, stream>>>(pint.as<int>(), n, -1);
        k_fill_i32<<<nblk(n), kBlock, 0, stream>>>(flags.as<int>(), n, 0);
        k_karras<<<nblk(n - 1), kBlock, 0, stream>>>(sk, n, left.as<int>(), right.as<int>(), first.as<int>(),
                                                     last.as<int>(), pint.as<int>(), pleaf.as<int>());
        k_refit<<<nblk(n), kBlock, 0, stream>>>(sk, n, boxes.as<Box6>(), left.as<int>(), right.as<int>(), pint.as<int>(),
                                                pleaf.as<int>(), flags.as<int>(), ibox.as<Box6>());
        // emitted internal nodes: prefix sum of (count > kLeafMax)
        k_emit_flags<<<nblk(n - 1), kBlock, 0, stream>>>(first.as<int>(), last.as<int>(), n - 1, flags.as<int>());
        tmp_bytes = 0;
        RTG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, flags.as<int>(), newidx.as<int>(), 0, (size_t)(n - 1),
                                        rocprim::plus<int>(), stream));
        DevBuf scan_tmp;
        if (ensure(scan_tmp, std::max<size_t>(tmp_bytes, 16))) return fail(RTG_E_NOMEM, "scan temp");
        RTG_HIP(rocprim::exclusive_scan(scan_tmp.p, tmp_bytes, flags.as<int>(), newidx.as<int>(), 0, (size_t)(n - 1),
                                        rocprim::plus<int>(), stream));
        int last_new = 0, last_flag = 0;
        RTG_HIP(hipMemcpyAsync(&last_new, newidx.as<int>() + (n - 2), sizeof(int), hipMemcpyDeviceToHost, stream));
        RTG_HIP(hipMemcpyAsync(&last_flag, flags.as<int>() + (n - 2), sizeof(int), hipMemcpyDeviceToHost, stream));
        RTG_HIP(hipStreamSynchronize(stream));
        n_emit = last_new + last_flag;
        k_top_first<<<1, 1, 0, stream>>>(left.as<int>(), right.as<int>(), first.as<int>(), last.as<int>(), n - 1,
                                         newidx.as<int>());
        if (ensure(nodes2, std::max(n_emit, 1) * sizeof(BvhNode))) return fail(RTG_E_NOMEM, "bvh nodes");
        k_emit<<<nblk(n - 1), kBlock, 0, stream>>>(sk, n - 1, boxes.as<Box6>(), left.as<int>(), right.as<int>(),
                                                   first.as<int>(), last.as<int>(), ibox.as<Box6>(), newidx.as<int>(),
                                                   nodes2.as<BvhNode>());
        if (ensure(S.prims, (size_t)n * 3 * sizeof(float4))) return fail(RTG_E_NOMEM, "prims");
        k_reorder<<<nblk(n), kBlock, 0, stream>>>(sk, n, prims_u.as<float4>(), S.prims.as<float4>());
        root = 0;
        S.topo_n = n, S.topo_emit = n_emit;
    } else {
        // few primitives: one leaf, prims in input order
        if (ensure(S.prims, std::max(n, 1) * 3 * sizeof(float4)) || ensure(S.nodes, sizeof(Bvh4Node)))
            return fail(RTG_E_NOMEM, "prims");
        if (n > 0)
            RTG_HIP(hipMemcpyAsync(S.prims.p, prims_u.p, (size_t)n * 3 * sizeof(float4), hipMemcpyDeviceToDevice,
                                   stream));
        root = n > 0 ? make_leaf(0, n) : 0;
    }
    int n4 = 0;
    if (n_emit > 0 && (st = widen_tree(nodes2, n_emit, S, stream, &n4)) != RTG_OK) return st;
    RTG_HIP(hipEventRecord(e1, stream));
    RTG_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    RTG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S.build_ms = ms;
    S.n_nodes = n4;
    shading_state(S, d);

    SceneView &v = S.view;
    v = SceneView{};
    v.nodes = S.nodes.as<Bvh4Node>();
    v.n_nodes = n4;
    v.prims = S.prims.as<float4>();
    v.root = root;
    v.lights_on = S.lights_on;
    v.n_prims = n;
    v.objs = S.objs.as<DevObject>();
    v.n_objs = d.n_objects;
    v.n_brute = (int)brute.size();
    v.brute = S.brute.as<int>();
    v.parts = S.parts.as<DevPart>();
    v.tri_nrm = S.tri_nrm.as<float4>();
    v.tri_uv = S.tri_uv.as<float2>();
    v.tri_pos = S.tri_pos.as<float4>();
    v.tri_part = S.tri_part.as<int>();
    v.tri_mtl = S.tri_mtl.as<int>();
    v.tri_bins = S.tri_bins.as<uint8_t>();
    v.mtls = S.mtls.as<DevMaterial>();
    v.lights = S.lights.as<DevLight>();
    v.n_lights = d.n_lights;
    v.n_texs = d.n_textures;
    v.texs = S.texs.as<DevTexture>();
    v.texels = S.texels.as<uint8_t>();
    v.env = prep.env;
    v.cam = prep.cam;
    S.camera = d.camera;
    S.bytes = scene_bytes(S);
    return RTG_OK;
}

// Scene update that leaves the geometry unchanged (materials, lights,
// textures, environment light, camera): re-upload those tables only; the
// BVH, primitives and per-triangle data stay.
int scene_update_shading(SceneRes &S, const rtg_scene_desc &d, hipStream_t stream) {
    ScenePrep prep;
    const std::string err = prepare_scene(d, prep);
    if (!err.empty()) return fail(RTG_E_INVALID, err);
    int st;
    if ((st = upload(S.objs, prep.objs, stream)) || (st = upload(S.mtls, prep.mtls, stream)) ||
        (st = upload(S.lights, prep.lights, stream)) || (st = upload(S.texs, prep.texs, stream)) ||
        (st = upload(S.texels, prep.texels, stream)))
        return st;
    RTG_HIP(hipStreamSynchronize(stream));  // renders on other streams see the new tables
    shading_state(S, d);
    SceneView &v = S.view;
    v.lights_on = S.lights_on;
    v.objs = S.objs.as<DevObject>();
    v.mtls = S.mtls.as<DevMaterial>();
    v.lights = S.lights.as<DevLight>();
    v.n_lights = d.n_lights;
    v.n_texs = d.n_textures;
    v.texs = S.texs.as<DevTexture>();
    v.texels = S.texels.as<uint8_t>();
    v.env = prep.env;
    v.cam = prep.cam;
    S.camera = d.camera;
    S.bytes = scene_bytes(S);
    return RTG_OK;
}

}  // namespace rtg
