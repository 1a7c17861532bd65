// rt_core.hpp — exact-order arithmetic and primitive tests of the render path.
//
// Every function here is __host__ __device__: the kernels use them on gfx950
// and the CPU emulation harness of the test suite runs the very same code.
// All sources are compiled with -ffp-contract=off and without fast-math, so
// each + - * / sqrt rounds once, as the reference's SSE code does; FP32
// division and square root are the correctly rounded sequences.
//
// Operation orders (reference 3DElement.cpp):
//   dot    (x x' + y y') + (z z' + 0)     DPPS 0x71             (206-214)
//   cross  (y z' - z y', z x' - x z', x y' - y x')               (190-197)
//   normal v / sqrt(dot(v, v)) per lane, true division           (218-238)
// The reference's fourth (w) lane never reaches an x, y or z result, so the
// device keeps three lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RTG_HD __host__ __device__ __forceinline__

namespace rtg {

struct V3 {
    float x, y, z;
};

RTG_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RTG_HD float as_float(int i) { return __builtin_bit_cast(float, i); }
RTG_HD int as_int(float f) { return __builtin_bit_cast(int, f); }

// Deepest ray tree: levels 0..RTG_MAX_LEVEL (rtg.h).
#define RTG_MAX_DEPTH 16
RTG_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RTG_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RTG_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
RTG_HD V3 mixmul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RTG_HD float dot(V3 a, V3 b) {
    float xy = a.x * b.x + a.y * b.y;
    float zw = a.z * b.z + 0.0f;
    return xy + zw;
}
RTG_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RTG_HD V3 normalize(V3 a) {
    float s = sqrtf(dot(a, a));
    return V3{a.x / s, a.y / s, a.z / s};
}
// _mm_min_ps / _mm_max_ps lane semantics (second operand on NaN)
RTG_HD float sse_min(float a, float b) { return a < b ? a : b; }
RTG_HD float sse_max(float a, float b) { return a > b ? a : b; }
// std::min / std::max semantics
RTG_HD float std_min(float a, float b) { return b < a ? b : a; }
RTG_HD float std_max(float a, float b) { return a < b ? b : a; }
// `fabs(x) < 1e-6` with the double constant of the reference: x <= 1e-6f
RTG_HD bool below_1e6(float x) { return fabsf(x) <= 9.99999997e-07f; }

constexpr float kMiss = 1e20f;

struct Ray {
    V3 o, d;
};

// BorderTest (Basic3DObject.cpp:44-81): entry distance of the ray into
// [mn, mx] clamped at 0, or 1e20 on a miss.
RTG_HD float border_test(const Ray &r, V3 mn, V3 mx) {
    V3 a = mn - r.o, b = mx - r.o;
    V3 rrd = V3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    a = mixmul(a, rrd);
    b = mixmul(b, rrd);
    float nx = sse_min(a.x, b.x), ny = sse_min(a.y, b.y), nz = sse_min(a.z, b.z);
    float fx = sse_max(a.x, b.x), fy = sse_max(a.y, b.y), fz = sse_max(a.z, b.z);
    if (below_1e6(r.d.y)) {
        if (r.o.y > mx.y || r.o.y < mn.y) return kMiss;
        ny = -1, fy = 1e10f;
    }
    if (below_1e6(r.d.x)) {
        if (r.o.x > mx.x || r.o.x < mn.x) return kMiss;
        nx = -1, fx = 1e10f;
    }
    if (below_1e6(r.d.z)) {
        if (r.o.z > mx.z || r.o.z < mn.z) return kMiss;
        nz = -1, fz = 1e10f;
    }
    float dmin = std_max(std_max(nx, ny), std_max(nz, 0.0f));
    float dmax = std_min(std_min(fx, fy), fz);
    if (dmax < dmin) return kMiss;
    return dmin;
}

// BorderTestEx (Model.cpp:482-664): slab test of the 8 half boxes of [mn, mx]
// split at their midpoint.  Lane b: bit 2 = upper x half, bit 1 = upper z
// half, bit 0 = upper y half.  Returns the smallest entry over hit lanes
// (1e20 if none) and the lane mask.
RTG_HD float border_test_ex(const Ray &r, V3 mn, V3 mx, uint32_t &mask) {
    mask = 0;
    const V3 mid = (mn + mx) * 0.5f;
    const V3 rrd = V3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    const V3 tmin = mixmul(mn - r.o, rrd), tmax = mixmul(mx - r.o, rrd), tmid = mixmul(mid - r.o, rrd);
    // per axis: low half [t(min), t(mid)], high half [t(mid), t(max)]
    float xl0 = sse_min(tmin.x, tmid.x), xl1 = sse_max(tmin.x, tmid.x);
    float xh0 = sse_min(tmid.x, tmax.x), xh1 = sse_max(tmid.x, tmax.x);
    float yl0 = sse_min(tmin.y, tmid.y), yl1 = sse_max(tmin.y, tmid.y);
    float yh0 = sse_min(tmid.y, tmax.y), yh1 = sse_max(tmid.y, tmax.y);
    float zl0 = sse_min(tmin.z, tmid.z), zl1 = sse_max(tmin.z, tmid.z);
    float zh0 = sse_min(tmid.z, tmax.z), zh1 = sse_max(tmid.z, tmax.z);
    if (below_1e6(r.d.y)) {
        if (r.o.y > mx.y || r.o.y < mn.y) return kMiss;
        if (r.o.y > mid.y)
            yl0 = kMiss, yl1 = 0, yh0 = 0, yh1 = kMiss;
        else if (r.o.y < mid.y)
            yl0 = 0, yl1 = kMiss, yh0 = kMiss, yh1 = 0;
        else
            yl0 = yh0 = 0, yl1 = yh1 = kMiss;
    }
    if (below_1e6(r.d.x)) {
        if (r.o.x > mx.x || r.o.x < mn.x) return kMiss;
        if (r.o.x > mid.x)
            xl0 = kMiss, xl1 = 0, xh0 = 0, xh1 = kMiss;
        else if (r.o.x < mid.x)
            xl0 = 0, xl1 = kMiss, xh0 = kMiss, xh1 = 0;
        else
            xl0 = xh0 = 0, xl1 = xh1 = kMiss;
    }
    if (below_1e6(r.d.z)) {
        if (r.o.z > mx.z || r.o.z < mn.z) return kMiss;
        if (r.o.z > mid.z)
            zl0 = kMiss, zl1 = 0, zh0 = 0, zh1 = kMiss;
        else if (r.o.z < mid.z)
            zl0 = 0, zl1 = kMiss, zh0 = kMiss, zh1 = 0;
        else
            zl0 = zh0 = 0, zl1 = zh1 = kMiss;
    }
    float minist = kMiss;
    for (int b = 0; b < 8; ++b) {
        float x0 = (b & 4) ? xh0 : xl0, x1 = (b & 4) ? xh1 : xl1;
        float y0 = (b & 1) ? yh0 : yl0, y1 = (b & 1) ? yh1 : yl1;
        float z0 = (b & 2) ? zh0 : zl0, z1 = (b & 2) ? zh1 : zl1;
        float amin = sse_max(x0, y0), amax = sse_min(x1, y1);
        amin = sse_max(amin, sse_max(z0, 0.0f));
        amax = sse_min(amax, z1);
        if (amin <= amax) {
            mask |= 1u << b;
            minist = std_min(minist, amin);
        }
    }
    return minist;
}

// TriangleTest (Model.cpp:714-744): Moller-Trumbore with a reciprocal of the
// determinant; accepts t > 1e-5 with barycentrics (w, u, v).
RTG_HD float tri_test(const Ray &r, V3 e1, V3 e2, V3 p0, float &w, float &u, float &v) {
    V3 tmp1 = cross(r.d, e2);
    V3 t2r = r.o - p0;
    float f = dot(e1, tmp1);
    f = 1.0f / f;
    u = dot(t2r, tmp1) * f;
    if (u < 0.0f || u > 1.0f) return kMiss;
    V3 tmp2 = cross(t2r, e1);
    v = dot(r.d, tmp2) * f;
    w = 1.0f - u - v;
    if (v < 0.0f || w < 0.0f) return kMiss;
    float t = dot(e2, tmp2) * f;
    return t > 1e-5f ? t : kMiss;
}

// Sphere outside path (Basic3DObject.cpp:170-177); 1e20 when rejected before t.
RTG_HD float sphere_enter_t(const Ray &r, V3 c, float rsq) {
    V3 s2r = r.o - c;
    float b = dot(r.d, s2r);
    if (b > 0) return kMiss;
    float dis = b * b - dot(s2r, s2r) + rsq;
    if (dis < 0) return kMiss;
    return -(b + sqrtf(dis));
}
// Sphere inside path: exit distance (Basic3DObject.cpp:145-148).
RTG_HD float sphere_exit_t(const Ray &r, V3 c, float rsq) {
    V3 s2r = r.o - c;
    float b = dot(r.d, s2r);
    float dis = b * b - dot(s2r, s2r) + rsq;
    return -b + sqrtf(dis);
}

// Plane distance (Basic3DObject.cpp:382-388); 1e20 when parallel or behind.
RTG_HD float plane_t(const Ray &r, V3 p0, V3 n) {
    float a = dot(r.d, n);
    if (below_1e6(a)) return kMiss;
    V3 p2r = r.o - p0;
    float b = dot(p2r, n);
    float dis = -b / a;
    if (dis < 0) return kMiss;
    return dis;
}
// Plane texture coordinates (Basic3DObject.cpp:392-396).
RTG_HD void plane_uv(const Ray &r, V3 p0, V3 ax, V3 ay, float &u, float &v) {
    V3 p2r = r.o - p0;
    V3 tmp1 = cross(r.d, ay);
    float f = 1.0f / dot(ax, tmp1) / 5.0f;
    u = dot(p2r, tmp1) * f;
    V3 tmp2 = cross(p2r, ax);
    v = dot(r.d, tmp2) * f;
}

// Color::put channel (3DElement.cpp:463-468); NaN quantises to 0.
RTG_HD uint8_t quantize(float c) {
    if (c > 1.0f) return 255;
    if (c < 0.0f) return 0;
    float s = c * 255.0f;
    return s == s ? (uint8_t)s : 0;
}

RTG_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace rtg
