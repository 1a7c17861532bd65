// rt_bvh.hpp — per-element steps of the device LBVH build.
//
// Build (one kernel per step, rtg_bvh.hip):
//   1. prim_setup      triangle -> clTri values (p0 + position, p1 - p0,
//                      p2 - p0, Model.cpp:426), octant-bin membership
//                      (Model.cpp:430-465), shading arrays, and an AABB of
//                      the implied vertices computed in FP64 and rounded
//                      outward to FP32, then padded by 2^-20 of its magnitude
//                      (covers the FP32 Moller-Trumbore acceptance fuzz).
//   2. morton          30-bit Morton code of the box centre; key = code<<32 | id
//                      (unique keys).
//   3. radix sort of the keys.
//   4. karras_node     binary radix tree (Karras 2012) over the sorted keys.
//   5. refit           bottom-up box union (exact: min/max).
//   6. emit            subtrees of <= kLeafMax primitives become leaves; the
//                      rest are written as 64-byte binary BvhNodes.
//   7. widen           every binary node at even depth becomes a 4-wide node
//                      (Bvh4Node) whose children are its children, each
//                      internal one replaced by its own two children; child
//                      boxes are quantised outward to 8 bits per bound.
#pragma once
#include "rt_scene.hpp"
#include "../../include/rtg.h"

namespace rtg {

constexpr int kLeafMax = 2;  // primitives per leaf at most

struct Box6 {
    float lo[3], hi[3];
};

RTG_HD Box6 box_union(const Box6 &a, const Box6 &b) {
    Box6 r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = fminf(a.lo[k], b.lo[k]);
        r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return r;
}

RTG_HD float round_down(double x) {
    float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}
RTG_HD float round_up(double x) {
    float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}

// FP64 box of points, rounded outward to FP32 and padded.
RTG_HD Box6 padded_box(const double lo[3], const double hi[3]) {
    Box6 b;
    float m = 1.0f;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = round_down(lo[k]);
        b.hi[k] = round_up(hi[k]);
        m = fmaxf(m, fmaxf(fabsf(b.lo[k]), fabsf(b.hi[k])));
    }
    const float pad = m * 9.5367431640625e-07f;  // 2^-20
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = b.lo[k] - pad;
        b.hi[k] = b.hi[k] + pad;
    }
    return b;
}

// Per-triangle setup.  `tri` is the model-space reference Triangle,
// `pos` the model position, `mid` the part's model-space border midpoint
// ((borders[2p] + borders[2p+1]) * 0.5, Model.cpp:420).
RTG_HD void prim_setup_tri(const rtg_triangle &tri, V3 pos, V3 mid, int g, int obj, float4 *prim3, float4 *nrm3,
                           float2 *uv3, float4 *pos3, uint8_t &bins, Box6 &box) {
    const V3 p0 = V3{tri.p[0].x, tri.p[0].y, tri.p[0].z}, p1 = V3{tri.p[1].x, tri.p[1].y, tri.p[1].z},
             p2 = V3{tri.p[2].x, tri.p[2].y, tri.p[2].z};
    const V3 e1 = p1 - p0, e2 = p2 - p0, pw = p0 + pos;
    prim3[0] = make_float4(pw.x, pw.y, pw.z, as_float(g));
    prim3[1] = make_float4(e1.x, e1.y, e1.z, as_float(obj));
    prim3[2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    for (int k = 0; k < 3; ++k) {
        nrm3[k] = make_float4(tri.n[k].x, tri.n[k].y, tri.n[k].z, 0.0f);
        uv3[k] = make_float2(tri.uv[2 * k], tri.uv[2 * k + 1]);
        pos3[k] = make_float4(tri.p[k].x, tri.p[k].y, tri.p[k].z, 0.0f);
    }
    // octant bins by the model-space triangle bounds (_mm_min_ps/_mm_max_ps)
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        const float a0 = (&tri.p[0].x)[k], a1 = (&tri.p[1].x)[k], a2 = (&tri.p[2].x)[k];
        const float m12 = sse_min(a1, a2), M12 = sse_max(a1, a2);
        mn[k] = sse_min(a0, m12);
        mx[k] = sse_max(a0, M12);
    }
    const float md[3] = {mid.x, mid.y, mid.z};
    uint32_t in = 0;
    if (mn[0] <= md[0]) {
        if (mn[2] <= md[2]) in |= (mn[1] <= md[1] ? 1u : 0u) | (mx[1] >= md[1] ? 2u : 0u);
        if (mx[2] >= md[2]) in |= (mn[1] <= md[1] ? 4u : 0u) | (mx[1] >= md[1] ? 8u : 0u);
    }
    if (mx[0] >= md[0]) {
        if (mn[2] <= md[2]) in |= (mn[1] <= md[1] ? 16u : 0u) | (mx[1] >= md[1] ? 32u : 0u);
        if (mx[2] >= md[2]) in |= (mn[1] <= md[1] ? 64u : 0u) | (mx[1] >= md[1] ? 128u : 0u);
    }
    bins = (uint8_t)in;
    // implied world vertices p0w, p0w + e1, p0w + e2, exact in FP64
    double lo[3], hi[3];
    const double w0[3] = {pw.x, pw.y, pw.z}, d1[3] = {e1.x, e1.y, e1.z}, d2[3] = {e2.x, e2.y, e2.z};
    for (int k = 0; k < 3; ++k) {
        const double a = w0[k], b = w0[k] + d1[k], c = w0[k] + d2[k];
        lo[k] = fmin(a, fmin(b, c));
        hi[k] = fmax(a, fmax(b, c));
    }
    box = padded_box(lo, hi);
}

RTG_HD void prim_setup_sphere(V3 c, float radius, float rsq, int obj, float4 *prim3, Box6 &box) {
    prim3[0] = make_float4(c.x, c.y, c.z, as_float(~obj));
    prim3[1] = make_float4(rsq, radius, 0.0f, 0.0f);
    prim3[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const double r = fabs((double)radius);
    const double lo[3] = {(double)c.x - r, (double)c.y - r, (double)c.z - r};
    const double hi[3] = {(double)c.x + r, (double)c.y + r, (double)c.z + r};
    box = padded_box(lo, hi);
}

RTG_HD uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton grid of the centroid bounds cb (lo x, y, z, hi x, y, z): origin and
// scale.  The cells are cubes, the longest axis split into 1024: a grid
// stretched to the bounds on every axis would split a flat scene (a height
// field) across its thin axis as often as along the others, and its tree
// visits 25% more nodes (CPU count, height field of workload 3).
RTG_HD void morton_grid(const float cb[6], float cmin[3], float cinv[3]) {
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) m = fmaxf(m, cb[3 + k] - cb[k]);
    for (int k = 0; k < 3; ++k) {
        cmin[k] = cb[k];
        cinv[k] = m > 0 ? 1.0f / m : 0.0f;
    }
}

RTG_HD uint64_t morton_key(const Box6 &b, const float cmin[3], const float cinv[3], int id) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float c = 0.5f * (b.lo[k] + b.hi[k]);
        float x = (c - cmin[k]) * cinv[k] * 1024.0f;
        x = fminf(fmaxf(x, 0.0f), 1023.0f);
        q[k] = (uint32_t)x;
    }
    const uint64_t code = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    return (code << 32) | (uint32_t)id;
}

RTG_HD int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

RTG_HD int delta(const uint64_t *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return clz64(keys[i] ^ keys[j]);
}

// Karras 2012, internal node i of n-1.  Children: >= 0 internal node index,
// < 0 leaf ~k (sorted position k).  Writes the node's key range [first, last].
RTG_HD void karras_node(const uint64_t *keys, int n, int i, int &left, int &right, int &first, int &last) {
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    first = i < j ? i : j;
    last = i < j ? j : i;
    left = first == gamma ? ~gamma : gamma;
    right = last == gamma + 1 ? ~(gamma + 1) : gamma + 1;
}

// Child link and box of a radix-tree child for the emitted layout.
RTG_HD int emit_link(int child, const int *first, const int *last, const int *newidx) {
    if (child < 0) return make_leaf(~child, 1);
    const int cnt = last[child] - first[child] + 1;
    if (cnt <= kLeafMax) return make_leaf(first[child], cnt);
    return newidx[child];
}

// Lay the top of the tree out first: the emitted nodes of the first BFS
// levels from the root (at most kTopNodes) take emitted indices 0 ..
// count - 1, in BFS order; the nodes that held those indices take the ones
// the top nodes leave, so newidx stays a permutation of the emitted nodes.
// Runs between the prefix sum of emitted nodes (newidx, non-decreasing over
// the n_int radix-tree nodes) and the emit; needs more than kLeafMax
// primitives (the root is emitted).  Returns the number of top nodes.
RTG_HD int top_first(const int *left, const int *right, const int *first, const int *last, int n_int, int *newidx) {
    auto emitted = [&](int v) { return v >= 0 && last[v] - first[v] + 1 > kLeafMax; };
    int q[kTopNodes];
    int nq = 0;
    q[nq++] = 0;
    for (int head = 0; head < nq && nq < kTopNodes; ++head) {
        const int c0 = left[q[head]], c1 = right[q[head]];
        if (emitted(c0) && nq < kTopNodes) q[nq++] = c0;
        if (emitted(c1) && nq < kTopNodes) q[nq++] = c1;
    }
    // holder[j]: the emitted node whose emitted index is j (the first
    // emitted v with newidx[v] == j)
    int holder[kTopNodes], cur[kTopNodes];
    for (int j = 0; j < nq; ++j) {
        int lo = 0, hi = n_int - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (newidx[mid] < j)
                lo = mid + 1;
            else
                hi = mid;
        }
        while (lo < n_int - 1 && !emitted(lo)) ++lo;
        holder[j] = lo;
        cur[j] = newidx[q[j]];
    }
    // each displaced holder takes an index a top node leaves (one outside
    // 0 .. nq - 1 for every holder that is not a top node)
    int k = 0;
    for (int j = 0; j < nq; ++j) {
        bool top = false;
        for (int i = 0; i < nq; ++i) top = top || holder[j] == q[i];
        if (top) continue;
        while (k < nq && cur[k] < nq) ++k;
        if (k >= nq) break;
        newidx[holder[j]] = cur[k++];
    }
    for (int j = 0; j < nq; ++j) newidx[q[j]] = j;
    return nq;
}

// ---- 4-wide nodes ---------------------------------------------------------

// The binary node's child `side` (0, 1): its box and link.
RTG_HD void node2_child(const BvhNode &n, int side, Box6 &b, int &link) {
    const float4 xy = side ? n.xy1 : n.xy0;
    b.lo[0] = xy.x, b.hi[0] = xy.y, b.lo[1] = xy.z, b.hi[1] = xy.w;
    b.lo[2] = side ? n.z01.z : n.z01.x, b.hi[2] = side ? n.z01.w : n.z01.y;
    link = side ? n.link.y : n.link.x;
}

// Children of the 4-wide node made from binary node i: each leaf child kept,
// each internal child replaced by its two children.  Links are in binary
// terms (leaf codes, or binary node indices at even depth).  Returns 2..4.
RTG_HD int node4_children(const BvhNode *nodes2, int i, Box6 box[4], int link[4]) {
    int c = 0;
    for (int side = 0; side < 2; ++side) {
        Box6 b;
        int l;
        node2_child(nodes2[i], side, b, l);
        if (l < 0) {
            box[c] = b, link[c] = l, ++c;
        } else {
            for (int s2 = 0; s2 < 2; ++s2) {
                node2_child(nodes2[l], s2, box[c], link[c]);
                ++c;
            }
        }
    }
    return c;
}

RTG_HD double box_area(const Box6 &b) {
    const double x = (double)b.hi[0] - b.lo[0], y = (double)b.hi[1] - b.lo[1], z = (double)b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

// Depth parity of binary node i (root 0 at depth 0) by its parent links.
RTG_HD bool node2_even(const int *parent2, int i) {
    int d = 0;
    while (i > 0 && d < 256) i = parent2[i], ++d;
    return (d & 1) == 0;
}

// Byte bounds of `nc` child boxes on one axis: the smallest power-of-two step
// s = 2^(e - 127) for which every child's [lo, hi] is covered by
// [org + qlo s, org + qhi s] with q in 0..255 and org the union's lower bound
// rounded down to a multiple m s of the step, |m| + 255 < 2^24: then every
// decoded bound (m + q) s is a float exactly, which the walk's child test
// relies on (Node4Test).  The covering is checked with the decode's own float
// arithmetic.  An axis no step covers (infinite bounds) gets org = -FLT_MAX,
// s = 2^127 and the full byte range, which decodes to [-FLT_MAX, +inf].
RTG_HD void quantize_axis(const Box6 *b, int nc, int a, float &org, int &e, uint8_t *qlo, uint8_t *qhi) {
    float lo = b[0].lo[a], hi = b[0].hi[a];
    for (int c = 1; c < nc; ++c) lo = fminf(lo, b[c].lo[a]), hi = fmaxf(hi, b[c].hi[a]);
    const double ext = (double)hi - (double)lo;
    int k = -126;  // step 2^k
    if (ext > 0.0 && ext < 1e300) {
        int x;
        const double m = frexp(ext / 255.0, &x);  // ext / 255 = m 2^x, m in [0.5, 1)
        k = m == 0.5 ? x - 1 : x;
        k = k < -126 ? -126 : k;
    }
    for (; k <= 127; ++k) {
        const float s = pow2f(k + 127);
        const double mo = floor((double)lo / (double)s);  // exact: lo / s is a power-of-two scaling
        if (!(fabs(mo) + 255.0 < 16777216.0)) continue;
        const float o = (float)(mo * (double)s);  // m s, exact
        bool ok = true;
        for (int c = 0; c < nc && ok; ++c) {
            const double fl = floor(((double)b[c].lo[a] - (double)o) / (double)s);
            const double fh = ceil(((double)b[c].hi[a] - (double)o) / (double)s);
            int ql = fl < 0 ? 0 : (fl > 255 ? 255 : (int)fl);
            int qh = fh < 0 ? 0 : (fh > 255 ? 255 : (int)fh);
            while (ql > 0 && !(dequant(o, (uint32_t)ql, s) <= b[c].lo[a])) --ql;
            while (qh < 255 && !(dequant(o, (uint32_t)qh, s) >= b[c].hi[a])) ++qh;
            ok = dequant(o, (uint32_t)ql, s) <= b[c].lo[a] && dequant(o, (uint32_t)qh, s) >= b[c].hi[a];
            qlo[c] = (uint8_t)ql, qhi[c] = (uint8_t)qh;
        }
        if (ok) {
            org = o, e = k + 127;
            return;
        }
    }
    org = -3.40282347e38f, e = 254;
    for (int c = 0; c < nc; ++c) qlo[c] = 0, qhi[c] = 255;
}

// The 4-wide node over `nc` (2..4) children with final links `link`.
RTG_HD Bvh4Node make_node4(const Box6 *box, const int *link, int nc) {
    Bvh4Node n;
    uint8_t ql[3][4] = {}, qh[3][4] = {};
    float org[3];
    int e[3];
    for (int a = 0; a < 3; ++a) quantize_axis(box, nc, a, org[a], e[a], ql[a], qh[a]);
    auto pack = [](const uint8_t *q) {
        return (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24;
    };
    const uint32_t meta = (uint32_t)e[0] | (uint32_t)e[1] << 8 | (uint32_t)e[2] << 16 | (uint32_t)nc << 24;
    n.org = make_float4(org[0], org[1], org[2], __builtin_bit_cast(float, meta));
    n.qxy = make_uint4(pack(ql[0]), pack(qh[0]), pack(ql[1]), pack(qh[1]));
    n.qz = make_uint4(pack(ql[2]), pack(qh[2]), 0u, 0u);
    n.link = make_int4(nc > 0 ? link[0] : 0, nc > 1 ? link[1] : 0, nc > 2 ? link[2] : 0, nc > 3 ? link[3] : 0);
    return n;
}

// The 4-wide node over `nc` children with final links.  The any-hit walk
// takes the children in stored order (closest-hit walks sort them by
// distance), so the order is chosen to meet an occluder early: leaves first,
// the smaller leaves first, then internal children by surface area, the
// largest first.
RTG_HD Bvh4Node ordered_node4(Box6 box[4], int link[4], int nc) {
    double key[4];
    for (int c = 0; c < nc; ++c)
        key[c] = link[c] < 0 ? leaf_count(link[c]) - 8.0 : 1.0 / (1.0 + box_area(box[c]));
    for (int a = 1; a < nc; ++a)  // stable insertion sort, ascending key
        for (int b = a; b > 0 && key[b] < key[b - 1]; --b) {
            const double tk = key[b];
            key[b] = key[b - 1], key[b - 1] = tk;
            const Box6 tb = box[b];
            box[b] = box[b - 1], box[b - 1] = tb;
            const int tl = link[b];
            link[b] = link[b - 1], link[b - 1] = tl;
        }
    return make_node4(box, link, nc);
}

// The 4-wide node made from binary node i (at even depth); idx4 maps binary
// node indices at even depth to 4-wide indices.
RTG_HD Bvh4Node widen_node(const BvhNode *nodes2, const int *idx4, int i) {
    Box6 box[4];
    int link[4];
    const int nc = node4_children(nodes2, i, box, link);
    for (int c = 0; c < nc; ++c)
        if (link[c] >= 0) link[c] = idx4[link[c]];
    return ordered_node4(box, link, nc);
}

RTG_HD BvhNode make_node(const Box6 &b0, const Box6 &b1, int l0, int l1) {
    BvhNode n;
    n.xy0 = make_float4(b0.lo[0], b0.hi[0], b0.lo[1], b0.hi[1]);
    n.xy1 = make_float4(b1.lo[0], b1.hi[0], b1.lo[1], b1.hi[1]);
    n.z01 = make_float4(b0.lo[2], b0.hi[2], b1.lo[2], b1.hi[2]);
    n.link = make_int4(l0, l1, 0, 0);
    return n;
}

}  // namespace rtg
