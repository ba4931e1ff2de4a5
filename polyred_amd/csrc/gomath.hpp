// gomath.hpp — polyred's float32 arithmetic, bit for bit, on host and device.
//
// The reference renderer is Go; its results depend on the exact sequence of
// float32 roundings.  These helpers reproduce that sequence:
//   * math.FMA[float32](x,y,z) = float32(math.FMA(float64 x,y,z))
//     (math/math.go:265-267): a double FMA (v_fma_f64 on gfx950) then one
//     rounding to float;
//   * Sqrt/Pow/Atan go through float64 (math/math.go:91-233): float32 sqrt of
//     a float32 is the correctly rounded f32 sqrt, which HIP emits by default;
//   * plain float32 expressions round once per operation (hipcc is always
//     invoked with -ffp-contract=off);
//   * generic Min/Max fold through float64 math.Min/Max from +-MaxFloat64
//     (math/math.go:245-261) with +Inf-before-NaN and +0-over--0 rules;
//   * float->int is Go's amd64 conversion (NaN/out of range -> MinInt64),
//     float->uint8 goes through int32 and keeps the low byte.
// Vectors/matrices follow math/vec3.go, math/vec4.go and math/mat4.go.
#pragma once

#include <stdint.h>
#include <float.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GM_HD __host__ __device__ __forceinline__
#else
#define GM_HD inline
#endif

#ifndef __HIP_DEVICE_COMPILE__
#include <cmath>
#endif

namespace gm {

GM_HD float fma64(float a, float b, float c) {
#ifdef __HIP_DEVICE_COMPILE__
    return (float)__builtin_fma((double)a, (double)b, (double)c);
#else
    return (float)std::fma((double)a, (double)b, (double)c);
#endif
}
GM_HD float sqrt32(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_sqrtf(x);
#else
    return (float)std::sqrt((double)x);
#endif
}
GM_HD float round32(float x) {  // math.Round: half away from zero
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_roundf(x);
#else
    return (float)std::round((double)x);
#endif
}
GM_HD double pow64(double x, double y) {
#ifdef __HIP_DEVICE_COMPILE__
    return pow(x, y);
#else
    return std::pow(x, y);
#endif
}
// float32(pow(x, y)) for the shading's x in (0, 1] and y in [2^-8, 2048]
// (the Blinn-Phong specular term), without the library's double-double pow:
// t = y log2(x) split as k + f (k an integer, |f| <= 1/2), log2 of the
// mantissa from the atanh series and 2^f from the exponential series, both
// in double to a relative error below 2^-40 of the result (the series are
// truncated below 2^-52; y e is exact; the other roundings are 2^-53 each).
// That decides the float32 rounding unless the value lies within 2^-40 of a
// float32 rounding midpoint; then -- and outside the ranges above -- the
// library pow decides, so the result is always (float)pow64(x, y).
GM_HD float pow32_fast(float xf, float yf) {
    const double x = xf, y = yf;
    // pow(+0, y) = +0 for every y > 0 (the specular term of a light behind
    // the surface's half vector), without the library
    uint32_t xb;
    __builtin_memcpy(&xb, &xf, 4);
    if (xb == 0u && yf > 0.0f) return 0.0f;
    if (!(xf > 0.0f && xf <= 1.0f && yf >= 0.00390625f && yf <= 2048.0f)) return (float)pow64(x, y);
    // x = m 2^e, m in [sqrt(1/2), sqrt(2))
    int e;
#ifdef __HIP_DEVICE_COMPILE__
    double m = __builtin_frexp(x, &e);  // [1/2, 1)
#else
    double m = std::frexp(x, &e);
#endif
    if (m < 0.70710678118654752) m *= 2.0, e -= 1;
    // log2(m) = (2 / ln 2) atanh(s), s = (m - 1) / (m + 1), |s| < 0.1716
    const double d = m + 1.0;
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(d);  // refined twice: relative error < 2^-56
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
#else
    const double r = 1.0 / d;
#endif
    const double sv = (m - 1.0) * r, s2 = sv * sv;
    // 2/ln2 * (1 + s^2/3 + s^4/5 + ... + s^18/19)
    double q = 2.0 / (19.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (17.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (15.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (13.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (11.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (9.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (7.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (5.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / (3.0 * 0.69314718055994531);
    q = q * s2 + 2.0 / 0.69314718055994531;
    const double l = sv * q;          // log2(m), |l| <= 1/2
    const double a = y * (double)e;   // exact: 24 x 9 bits
    const double b = y * l;
    const double t = a + b;
    if (t < -152.0) return 0.0f;       // below 2^-152: rounds to +0
#ifdef __HIP_DEVICE_COMPILE__
    const double kd = __builtin_rint(t);
#else
    const double kd = std::nearbyint(t);
#endif
    const double f = (a - kd) + b;    // a - kd exact; |f| <= 1/2 + 2^-40
    // 2^f = sum (f ln2)^n / n!, n <= 12
    const double z = f * 0.69314718055994531;
    double p = 1.0 / 479001600.0;
    p = p * z + 1.0 / 39916800.0;
    p = p * z + 1.0 / 3628800.0;
    p = p * z + 1.0 / 362880.0;
    p = p * z + 1.0 / 40320.0;
    p = p * z + 1.0 / 5040.0;
    p = p * z + 1.0 / 720.0;
    p = p * z + 1.0 / 120.0;
    p = p * z + 1.0 / 24.0;
    p = p * z + 1.0 / 6.0;
    p = p * z + 0.5;
    p = p * z + 1.0;
    p = p * z + 1.0;
#ifdef __HIP_DEVICE_COMPILE__
    const double v = __builtin_ldexp(p, (int)kd);
#else
    const double v = std::ldexp(p, (int)kd);
#endif
    // the float32 rounding of v holds for the exact value when v is farther
    // than the error bound from both neighbouring rounding midpoints
    const float fv = (float)v;
    uint32_t bits;
    __builtin_memcpy(&bits, &fv, 4);
    float up, dn;
    const uint32_t ub = bits + 1u, db = bits == 0u ? 0x80000001u : bits - 1u;
    __builtin_memcpy(&up, &ub, 4);
    __builtin_memcpy(&dn, &db, 4);
    const double mid_up = 0.5 * ((double)fv + (double)up), mid_dn = 0.5 * ((double)fv + (double)dn);
    const double margin = v * 0x1p-40;
    if (mid_up - v > margin && v - mid_dn > margin) return fv;
    return (float)pow64(x, y);
}
GM_HD double atan64(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return atan(x);
#else
    return std::atan(x);
#endif
}
GM_HD bool isnan32(float x) { return x != x; }
GM_HD bool isinf64(double x) { return x - x != 0 && x == x; }
GM_HD bool signbit64(double x) {
    union { double d; uint64_t u; } c;
    c.d = x;
    return (c.u >> 63) != 0;
}
GM_HD double go_max64(double x, double y) {
    if ((isinf64(x) && x > 0) || (isinf64(y) && y > 0)) return __builtin_inf();
    if (x != x || y != y) return __builtin_nan("");
    if (x == 0 && x == y) return signbit64(x) ? y : x;
    return x > y ? x : y;
}
GM_HD double go_min64(double x, double y) {
    if ((isinf64(x) && x < 0) || (isinf64(y) && y < 0)) return -__builtin_inf();
    if (x != x || y != y) return __builtin_nan("");
    if (x == 0 && x == y) return signbit64(x) ? x : y;
    return x < y ? x : y;
}
// generic Max[float32](a,b) / Min[float32](a,b)
GM_HD float max2(float a, float b) { return (float)go_max64(go_max64(-DBL_MAX, a), b); }
GM_HD float min2(float a, float b) { return (float)go_min64(go_min64(DBL_MAX, a), b); }

GM_HD int64_t f2i(float x) {
    if (x >= -9223372036854775808.0f && x < 9223372036854775808.0f) return (int64_t)x;
    return INT64_MIN;
}
GM_HD uint8_t f2u8(float x) {
    if (x >= -2147483648.0f && x < 2147483648.0f) return (uint8_t)(uint32_t)(int32_t)x;
    return 0;
}
GM_HD float clamp(float n, float lo, float hi) {  // math/clamp.go:8-16
    if (n < lo) return lo;
    if (n > hi) return hi;
    return n;
}
GM_HD float abs32(float x) { return x < 0 ? -x : (x == 0 ? 0.0f : x); }
GM_HD bool approx_eq(float a, float b, float eps) { return abs32(a - b) <= eps; }

struct V4 {
    float x, y, z, w;
};
struct V3 {
    float x, y, z;
};
GM_HD V4 v4(float x, float y, float z, float w) { return V4{x, y, z, w}; }
GM_HD V4 sub(V4 a, V4 b) { return v4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
GM_HD V4 add(V4 a, V4 b) { return v4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
GM_HD V4 scale(V4 a, float x, float y, float z, float w) { return v4(a.x * x, a.y * y, a.z * z, a.w * w); }
GM_HD float dot(V4 v, V4 u) { return fma64(v.x, u.x, fma64(v.y, u.y, fma64(v.z, u.z, v.w * u.w))); }
GM_HD float len(V4 v) { return sqrt32(dot(v, v)); }
GM_HD V4 unit(V4 v) {
    float n = 1.0f / len(v);
    return v4(v.x * n, v.y * n, v.z * n, v.w * n);
}
GM_HD V4 cross(V4 v, V4 u) {
    return v4(fma64(v.y, u.z, -v.z * u.y), fma64(v.z, u.x, -v.x * u.z), fma64(v.x, u.y, -v.y * u.x), 0);
}
GM_HD V4 pos(V4 v) {
    if (v.w == 1 || v.w == 0) return v4(v.x, v.y, v.z, 1);
    float inv = 1.0f / v.w;
    return v4(v.x * inv, v.y * inv, v.z * inv, 1);
}
GM_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
GM_HD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
GM_HD float dot(V3 v, V3 u) { return fma64(v.x, u.x, fma64(v.y, u.y, v.z * u.z)); }
GM_HD V3 unit(V3 v) {
    float n = 1.0f / sqrt32(dot(v, v));
    return v3(v.x * n, v.y * n, v.z * n);
}
GM_HD V3 cross(V3 v, V3 u) {
    return v3(fma64(v.y, u.z, -v.z * u.y), fma64(v.z, u.x, -v.x * u.z), fma64(v.x, u.y, -v.y * u.x));
}

// Row-major 4x4 (X00..X33).
struct M4 {
    float a[16];
};
GM_HD V4 apply(V4 v, const M4& m) {  // math/vec4.go:107-116
    const float* a = m.a;
    return v4(fma64(a[0], v.x, fma64(a[1], v.y, fma64(a[2], v.z, a[3] * v.w))),
              fma64(a[4], v.x, fma64(a[5], v.y, fma64(a[6], v.z, a[7] * v.w))),
              fma64(a[8], v.x, fma64(a[9], v.y, fma64(a[10], v.z, a[11] * v.w))),
              fma64(a[12], v.x, fma64(a[13], v.y, fma64(a[14], v.z, a[15] * v.w))));
}
GM_HD V4 mulv(const M4& m, V4 v) {  // math/mat4.go:224-231
    const float* a = m.a;
    return v4(a[0] * v.x + a[1] * v.y + a[2] * v.z + a[3] * v.w, a[4] * v.x + a[5] * v.y + a[6] * v.z + a[7] * v.w,
              a[8] * v.x + a[9] * v.y + a[10] * v.z + a[11] * v.w,
              a[12] * v.x + a[13] * v.y + a[14] * v.z + a[15] * v.w);
}

// Screen-space barycentric coordinates, math/interpolate.go:66-78.  The
// per-triangle part (edges ab, ac, bc and the doubled area Sabc) is split
// out so a rasteriser can hoist it; the result is bit-identical.
struct TriSetup {
    float ax, ay, bx, by;
    float abx, aby, acx, acy, bcx, bcy, sabc;
};
GM_HD TriSetup tri_setup(float ax, float ay, float bx, float by, float cx, float cy) {
    TriSetup t;
    t.ax = ax, t.ay = ay, t.bx = bx, t.by = by;
    t.abx = bx - ax, t.aby = by - ay;
    t.acx = cx - ax, t.acy = cy - ay;
    t.bcx = cx - bx, t.bcy = cy - by;
    t.sabc = fma64(t.abx, t.acy, -t.aby * t.acx);
    return t;
}
// The three edge areas (Sbcp, Sapc, Sabp) at p.
GM_HD void tri_areas(const TriSetup& t, float px, float py, float s[3]) {
    float apx = px - t.ax, apy = py - t.ay;
    float bpx = px - t.bx, bpy = py - t.by;
    s[2] = fma64(t.abx, apy, -t.aby * apx);  // Sabp
    s[1] = fma64(apx, t.acy, -apy * t.acx);  // Sapc
    s[0] = fma64(t.bcx, bpy, -t.bcy * bpx);  // Sbcp
}
GM_HD void barycoord(const TriSetup& t, float px, float py, float bc[3]) {
    float s[3];
    tri_areas(t, px, py, s);
    bc[0] = s[0] / t.sabc;
    bc[1] = s[1] / t.sabc;
    bc[2] = s[2] / t.sabc;
}
GM_HD bool inside(const float bc[3]) {  // render/raster.go:491
    return !(bc[0] < -1e-7f || bc[1] < -1e-7f || bc[2] < -1e-7f);
}
// The same inside test and the depth (render/raster.go:487-496) with the
// divisions only for covered points: fl(S / Sabc) < float32(-1e-7) exactly
// when S / Sabc < M, M the midpoint between float32(-1e-7) and the float32
// below it (a quotient of two float32 is never such a midpoint), and for a
// finite non-zero Sabc, S < M * Sabc (S > M * Sabc when Sabc < 0) decides that
// exactly, the product of M (25 bits) and Sabc (24 bits) being exact in
// float64.  Other Sabc take the divisions.
constexpr double BC_EDGE_MID = -0x1.ad7f2bp-24;
// IEEE minNum / maxNum: a NaN operand gives the other one
GM_HD float gm_fmin(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_fminf(a, b);
#else
    return std::fmin(a, b);
#endif
}
GM_HD float gm_fmax(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_fmaxf(a, b);
#else
    return std::fmax(a, b);
#endif
}
GM_HD bool cover_z(const TriSetup& t, float px, float py, float az, float bz, float cz, float& z) {
    float s[3];
    tri_areas(t, px, py, s);
    if (t.sabc != 0.0f && t.sabc - t.sabc == 0.0f) {  // finite, non-zero
        // (some S beyond M * Sabc: the least (greatest) of the three beyond
        // it, NaNs left out by fmin/fmax as by the comparisons)
        const double ms = BC_EDGE_MID * (double)t.sabc;
        const bool out = t.sabc > 0.0f ? (double)gm_fmin(gm_fmin(s[0], s[1]), s[2]) < ms
                                       : (double)gm_fmax(gm_fmax(s[0], s[1]), s[2]) > ms;
        if (out) return false;
    }
    const float bc[3] = {s[0] / t.sabc, s[1] / t.sabc, s[2] / t.sabc};
    if (!inside(bc)) return false;
    z = bc[0] * az + bc[1] * bz + bc[2] * cz;
    return true;
}

// Depth ordering for a 64-bit atomic-min visibility key: larger depth gives a
// smaller key; -0 and +0 tie (they compare equal under '>').  NaN is handled
// out of band by the caller.
GM_HD uint32_t depth_key_hi(float z) {
    union { float f; uint32_t u; } c;
    c.f = (z == 0.0f) ? 0.0f : z;
    uint32_t ord = (c.u & 0x80000000u) ? ~c.u : (c.u | 0x80000000u);
    return ~ord;
}
GM_HD float depth_from_key_hi(uint32_t hi) {
    uint32_t ord = ~hi;
    union { float f; uint32_t u; } c;
    c.u = (ord & 0x80000000u) ? (ord & 0x7fffffffu) : ~ord;
    return c.f;
}
// Ordered bits for atomic max of a shadow-map depth (value only).
GM_HD uint32_t depth_ord(float z) {
    union { float f; uint32_t u; } c;
    c.f = (z == 0.0f) ? 0.0f : z;
    return (c.u & 0x80000000u) ? ~c.u : (c.u | 0x80000000u);
}
GM_HD float depth_from_ord(uint32_t ord) {
    union { float f; uint32_t u; } c;
    c.u = (ord & 0x80000000u) ? (ord & 0x7fffffffu) : ~ord;
    return c.f;
}

}  // namespace gm
