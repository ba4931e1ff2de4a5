/*
 * oracle.c — TEST INFRASTRUCTURE ONLY (see oracle.h for the pinning list).
 *
 * Plain-C restatement of polyred's CPU rasteriser and deferred shader,
 * sequential (Workers(1)) order.  Every function names the reference lines it
 * follows.  Build flags matter: -O2 -ffp-contract=off -fno-fast-math on
 * x86-64 (SSE), so every float operation rounds once, as Go's amd64 compiler
 * (GOAMD64=v1) evaluates the reference.
 *
 * Go semantics reproduced here:
 *   - math.FMA[float32](x,y,z) = float32(math.FMA(float64 x,y,z))
 *     (math/math.go:265-267): a double FMA then a rounding to float;
 *   - Sqrt/Pow/Tan/Cos/Sin/Atan/Round go through float64 and back
 *     (math/math.go:58-233);
 *   - generic Min/Max fold through math.Min/math.Max on float64 starting at
 *     +-MaxFloat64 (math/math.go:245-261), with math.Max's special cases
 *     (+Inf first, then NaN, +0 preferred over -0);
 *   - float->int is CVTTSS2SQ (NaN and out-of-range give MinInt64) and
 *     float->uint8 goes through int32 then keeps the low byte.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } v4;
typedef struct { float x, y, z; } v3;
typedef struct { float a[16]; } m4; /* row-major, a[r*4+c] = Xrc */

/* ------------------------------------------------------------------ Go math */
static inline float gfma(float a, float b, float c) { return (float)fma((double)a, (double)b, (double)c); }
static inline float gsqrt(float x) { return (float)sqrt((double)x); }
static inline float ground(float x) { return (float)round((double)x); }
static inline float gpow(float x, float y) { return (float)pow((double)x, (double)y); }

static double go_max64(double x, double y) {
    if ((isinf(x) && x > 0) || (isinf(y) && y > 0)) return INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? y : x;
    return x > y ? x : y;
}
static double go_min64(double x, double y) {
    if ((isinf(x) && x < 0) || (isinf(y) && y < 0)) return -INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? x : y;
    return x < y ? x : y;
}
/* generic Max[float32](a, b) / Min[float32](a, b) */
static inline float gmax2(float a, float b) { return (float)go_max64(go_max64(-DBL_MAX, a), b); }
static inline float gmin2(float a, float b) { return (float)go_min64(go_min64(DBL_MAX, a), b); }

static inline int64_t gf2i(float x) {
    if (x >= -9223372036854775808.0f && x < 9223372036854775808.0f) return (int64_t)x;
    return INT64_MIN;
}
static inline uint8_t gf2u8(float x) {
    if (x >= -2147483648.0f && x < 2147483648.0f) return (uint8_t)(uint32_t)(int32_t)x;
    return 0;
}
static inline float gclamp(float n, float lo, float hi) { /* math/clamp.go:8-16 */
    if (n < lo) return lo;
    if (n > hi) return hi;
    return n;
}
static inline int approx_eq(float a, float b, float eps) { return (float)fabs((double)(a - b)) <= eps; }

/* ---------------------------------------------------------------- vectors */
static inline v4 V4(float x, float y, float z, float w) { v4 r = {x, y, z, w}; return r; }
static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v4 sub4(v4 a, v4 b) { return V4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline v4 add4(v4 a, v4 b) { return V4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4 scale4(v4 a, float x, float y, float z, float w) { return V4(a.x * x, a.y * y, a.z * z, a.w * w); }
/* math/vec4.go:89-91 */
static inline float dot4(v4 v, v4 u) { return gfma(v.x, u.x, gfma(v.y, u.y, gfma(v.z, u.z, v.w * u.w))); }
static inline float len4(v4 v) { return gsqrt(dot4(v, v)); }
/* math/vec4.go:96-104 */
static inline v4 unit4(v4 v) { float n = 1.0f / len4(v); return V4(v.x * n, v.y * n, v.z * n, v.w * n); }
/* math/vec4.go:107-116 */
static inline v4 apply4(v4 v, const m4* m) {
    const float* a = m->a;
    return V4(gfma(a[0], v.x, gfma(a[1], v.y, gfma(a[2], v.z, a[3] * v.w))),
              gfma(a[4], v.x, gfma(a[5], v.y, gfma(a[6], v.z, a[7] * v.w))),
              gfma(a[8], v.x, gfma(a[9], v.y, gfma(a[10], v.z, a[11] * v.w))),
              gfma(a[12], v.x, gfma(a[13], v.y, gfma(a[14], v.z, a[15] * v.w))));
}
/* math/vec4.go:127-137 */
static inline v4 cross4(v4 v, v4 u) {
    return V4(gfma(v.y, u.z, -v.z * u.y), gfma(v.z, u.x, -v.x * u.z), gfma(v.x, u.y, -v.y * u.x), 0);
}
/* math/vec4.go:140-146 */
static inline v4 pos4(v4 v) {
    if (v.w == 1 || v.w == 0) return V4(v.x, v.y, v.z, 1);
    float inv = 1.0f / v.w;
    return V4(v.x * inv, v.y * inv, v.z * inv, 1);
}
static inline v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
/* math/vec3.go:78-120 */
static inline float dot3(v3 v, v3 u) { return gfma(v.x, u.x, gfma(v.y, u.y, v.z * u.z)); }
static inline v3 unit3(v3 v) { float n = 1.0f / gsqrt(dot3(v, v)); return V3(v.x * n, v.y * n, v.z * n); }
static inline v3 cross3(v3 v, v3 u) {
    return V3(gfma(v.y, u.z, -v.z * u.y), gfma(v.z, u.x, -v.x * u.z), gfma(v.x, u.y, -v.y * u.x));
}

/* ---------------------------------------------------------------- matrices */
/* math/mat4.go:201-219: each entry a left-to-right sum of four products. */
static m4 mulm(const m4* m, const m4* n) {
    m4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.a[i * 4 + j] = m->a[i * 4 + 0] * n->a[0 * 4 + j] + m->a[i * 4 + 1] * n->a[1 * 4 + j] +
                             m->a[i * 4 + 2] * n->a[2 * 4 + j] + m->a[i * 4 + 3] * n->a[3 * 4 + j];
    return r;
}
/* math/mat4.go:224-231 */
static v4 mulv(const m4* m, v4 v) {
    const float* a = m->a;
    return V4(a[0] * v.x + a[1] * v.y + a[2] * v.z + a[3] * v.w, a[4] * v.x + a[5] * v.y + a[6] * v.z + a[7] * v.w,
              a[8] * v.x + a[9] * v.y + a[10] * v.z + a[11] * v.w,
              a[12] * v.x + a[13] * v.y + a[14] * v.z + a[15] * v.w);
}
static m4 transpose(const m4* m) {
    m4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.a[i * 4 + j] = m->a[j * 4 + i];
    return r;
}
/* Determinant terms in the order of math/mat4.go:233-247: {sign, i, j, k, l},
 * each term ((a_i*a_j)*a_k)*a_l, accumulated left to right. */
static const signed char DET_T[24][5] = {
    {+1, 0, 5, 10, 15}, {-1, 0, 5, 11, 14}, {+1, 0, 6, 11, 13}, {-1, 0, 6, 9, 15},  {+1, 0, 7, 9, 14},
    {-1, 0, 7, 10, 13}, {-1, 1, 6, 11, 12}, {+1, 1, 6, 8, 15},  {-1, 1, 7, 8, 14},  {+1, 1, 7, 10, 12},
    {-1, 1, 4, 10, 15}, {+1, 1, 4, 11, 14}, {+1, 2, 7, 8, 13},  {-1, 2, 7, 9, 12},  {+1, 2, 4, 9, 15},
    {-1, 2, 4, 11, 13}, {+1, 2, 5, 11, 12}, {-1, 2, 5, 8, 15},  {-1, 3, 4, 9, 14},  {+1, 3, 4, 10, 13},
    {-1, 3, 5, 10, 12}, {+1, 3, 5, 8, 14},  {-1, 3, 6, 8, 13},  {+1, 3, 6, 9, 12}};
/* Cofactor terms of math/mat4.go:259-282 per output entry: {sign, i, j, k}. */
static const signed char INV_T[16][6][4] = {
    {{+1, 6, 11, 13}, {-1, 7, 10, 13}, {+1, 7, 9, 14}, {-1, 5, 11, 14}, {-1, 6, 9, 15}, {+1, 5, 10, 15}},
    {{+1, 3, 10, 13}, {-1, 2, 11, 13}, {-1, 3, 9, 14}, {+1, 1, 11, 14}, {+1, 2, 9, 15}, {-1, 1, 10, 15}},
    {{+1, 2, 7, 13}, {-1, 3, 6, 13}, {+1, 3, 5, 14}, {-1, 1, 7, 14}, {-1, 2, 5, 15}, {+1, 1, 6, 15}},
    {{+1, 3, 6, 9}, {-1, 2, 7, 9}, {-1, 3, 5, 10}, {+1, 1, 7, 10}, {+1, 2, 5, 11}, {-1, 1, 6, 11}},
    {{+1, 7, 10, 12}, {-1, 6, 11, 12}, {-1, 7, 8, 14}, {+1, 4, 11, 14}, {+1, 6, 8, 15}, {-1, 4, 10, 15}},
    {{+1, 2, 11, 12}, {-1, 3, 10, 12}, {+1, 3, 8, 14}, {-1, 0, 11, 14}, {-1, 2, 8, 15}, {+1, 0, 10, 15}},
    {{+1, 3, 6, 12}, {-1, 2, 7, 12}, {-1, 3, 4, 14}, {+1, 0, 7, 14}, {+1, 2, 4, 15}, {-1, 0, 6, 15}},
    {{+1, 2, 7, 8}, {-1, 3, 6, 8}, {+1, 3, 4, 10}, {-1, 0, 7, 10}, {-1, 2, 4, 11}, {+1, 0, 6, 11}},
    {{+1, 5, 11, 12}, {-1, 7, 9, 12}, {+1, 7, 8, 13}, {-1, 4, 11, 13}, {-1, 5, 8, 15}, {+1, 4, 9, 15}},
    {{+1, 3, 9, 12}, {-1, 1, 11, 12}, {-1, 3, 8, 13}, {+1, 0, 11, 13}, {+1, 1, 8, 15}, {-1, 0, 9, 15}},
    {{+1, 1, 7, 12}, {-1, 3, 5, 12}, {+1, 3, 4, 13}, {-1, 0, 7, 13}, {-1, 1, 4, 15}, {+1, 0, 5, 15}},
    {{+1, 3, 5, 8}, {-1, 1, 7, 8}, {-1, 3, 4, 9}, {+1, 0, 7, 9}, {+1, 1, 4, 11}, {-1, 0, 5, 11}},
    {{+1, 6, 9, 12}, {-1, 5, 10, 12}, {-1, 6, 8, 13}, {+1, 4, 10, 13}, {+1, 5, 8, 14}, {-1, 4, 9, 14}},
    {{+1, 1, 10, 12}, {-1, 2, 9, 12}, {+1, 2, 8, 13}, {-1, 0, 10, 13}, {-1, 1, 8, 14}, {+1, 0, 9, 14}},
    {{+1, 2, 5, 12}, {-1, 1, 6, 12}, {-1, 2, 4, 13}, {+1, 0, 6, 13}, {+1, 1, 4, 14}, {-1, 0, 5, 14}},
    {{+1, 1, 6, 8}, {-1, 2, 5, 8}, {+1, 2, 4, 9}, {-1, 0, 6, 9}, {-1, 1, 4, 10}, {+1, 0, 5, 10}}};

static float det(const m4* m) {
    const float* a = m->a;
    float acc = 0;
    for (int t = 0; t < 24; t++) {
        float p = a[DET_T[t][1]] * a[DET_T[t][2]] * a[DET_T[t][3]] * a[DET_T[t][4]];
        acc = t == 0 ? p : (DET_T[t][0] > 0 ? acc + p : acc - p);
    }
    return acc;
}
/* math/mat4.go:255-283 (the zero-determinant panic becomes an identity-free
 * NaN result; the reference scenes never hit it). */
static m4 inv(const m4* m) {
    const float* a = m->a;
    float dinv = 1 / det(m);
    m4 r;
    for (int e = 0; e < 16; e++) {
        float acc = 0;
        for (int t = 0; t < 6; t++) {
            float p = a[INV_T[e][t][1]] * a[INV_T[e][t][2]] * a[INV_T[e][t][3]];
            acc = t == 0 ? p : (INV_T[e][t][0] > 0 ? acc + p : acc - p);
        }
        r.a[e] = dinv * acc;
    }
    return r;
}
/* math/math.go:270-277 */
static m4 viewport_matrix(float w, float h) {
    m4 r = {{w / 2, 0, 0, w / 2, 0, h / 2, 0, h / 2, 0, 0, 1, 0, 0, 0, 0, 1}};
    return r;
}
/* camera/camera.go:42-53 */
static m4 view_matrix(v3 pos, v3 target, v3 up) {
    v3 l = unit3(sub3(target, pos));
    v3 lxu = unit3(cross3(l, up));
    v3 u = unit3(cross3(lxu, l));
    m4 r = {{lxu.x, lxu.y, lxu.z, -dot3(lxu, pos), u.x, u.y, u.z, -dot3(u, pos), -l.x, -l.y, -l.z, dot3(l, pos), 0, 0,
             0, 1}};
    return r;
}
/* camera/perspective.go:100-111 */
static m4 persp_proj(float fov_deg, float aspect, float n, float f) {
    float fov = (fov_deg * (float)M_PI) / 180;
    float t = (float)tan((double)(fov / 2));
    m4 r = {{-1 / (aspect * t), 0, 0, 0, 0, -1 / t, 0, 0, 0, 0, (n + f) / (n - f), (2 * n * f) / (n - f), 0, 0, 1, 0}};
    return r;
}
/* camera/orthographic.go:106-117 */
static m4 ortho_proj(float l, float r_, float b, float t, float n, float f) {
    m4 r = {{2 / (r_ - l), 0, 0, (l + r_) / (l - r_), 0, 2 / (t - b), 0, (b + t) / (b - t), 0, 0, 2 / (n - f),
             (f + n) / (f - n), 0, 0, 0, 1}};
    return r;
}
static m4 cam_proj(const prh_camera* c) {
    if (c->kind == PRH_CAMERA_PERSPECTIVE) return persp_proj(c->fov, c->aspect, c->near_, c->far_);
    return ortho_proj(c->left, c->right, c->bottom, c->top, c->near_, c->far_);
}
static inline v3 V3a(const float* p) { return V3(p[0], p[1], p[2]); }

/* -------------------------------------------------------------- AABB etc. */
typedef struct { v3 mn, mx; } aabb_t;
/* geometry/primitive/box.go:15-25 */
static aabb_t aabb_new(int n, const v3* vs) {
    aabb_t b = {V3(FLT_MAX, FLT_MAX, FLT_MAX), V3(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
    for (int i = 0; i < n; i++) {
        b.mn.x = gmin2(b.mn.x, vs[i].x);
        b.mn.y = gmin2(b.mn.y, vs[i].y);
        b.mn.z = gmin2(b.mn.z, vs[i].z);
        b.mx.x = gmax2(b.mx.x, vs[i].x);
        b.mx.y = gmax2(b.mx.y, vs[i].y);
        b.mx.z = gmax2(b.mx.z, vs[i].z);
    }
    return b;
}
/* geometry/primitive/box.go:32-41, including maxZ = min(a.Max.Y, b.Max.Z) */
static int aabb_intersect(const aabb_t* a, const aabb_t* b) {
    float minX = gmax2(a->mn.x, b->mn.x), minY = gmax2(a->mn.y, b->mn.y), minZ = gmax2(a->mn.z, b->mn.z);
    float maxX = gmin2(a->mx.x, b->mx.x), maxY = gmin2(a->mx.y, b->mx.y), maxZ = gmin2(a->mx.y, b->mx.z);
    return minX <= maxX && minY <= maxY && minZ <= maxZ;
}
/* geometry/primitive/box.go:61-75 with math.ApproxEq/ApproxLess (math/math.go:41-49) */
static inline int less_eq(float v1, float v2) {
    const float eps = 1e-7f;
    return approx_eq(v1, v2, eps) || (v1 < v2 && (float)fabs((double)(v1 - v2)) > eps);
}
static int aabb_contains(const aabb_t* b, int n, const v3* vs) {
    for (int i = 0; i < n; i++) {
        int ok = less_eq(b->mn.x, vs[i].x) && less_eq(b->mn.y, vs[i].y) && less_eq(b->mn.z, vs[i].z) &&
                 less_eq(vs[i].x, b->mx.x) && less_eq(vs[i].y, b->mx.y) && less_eq(vs[i].z, b->mx.z);
        if (!ok) return 0;
    }
    return 1;
}
/* geometry/primitive/triangle.go:63-80 */
static int tri_is_valid(v4 p1, v4 p2, v4 p3) {
    const float eps = 1e-7f;
    v4 a = sub4(p2, p1), b = sub4(p3, p1);
    if (approx_eq(a.x, 0, eps) && approx_eq(a.y, 0, eps) && approx_eq(a.z, 0, eps)) return 0;
    if (approx_eq(b.x, 0, eps) && approx_eq(b.y, 0, eps) && approx_eq(b.z, 0, eps)) return 0;
    float d = dot4(a, b) / (len4(a) * len4(b));
    return !approx_eq(d, 1, eps) && !approx_eq(d, -1, eps);
}
/* math/interpolate.go:66-78 */
static inline void barycoord(float px, float py, float ax, float ay, float bx, float by, float cx, float cy, float out[3]) {
    v3 ap = V3(px - ax, py - ay, 0), ab = V3(bx - ax, by - ay, 0), ac = V3(cx - ax, cy - ay, 0);
    v3 bc = V3(cx - bx, cy - by, 0), bp = V3(px - bx, py - by, 0);
    float Sabc = cross3(ab, ac).z, Sabp = cross3(ab, ap).z, Sapc = cross3(ap, ac).z, Sbcp = cross3(bc, bp).z;
    out[0] = Sbcp / Sabc;
    out[1] = Sapc / Sabc;
    out[2] = Sabp / Sabc;
}

/* ------------------------------------------------------------ sRGB (color) */
static double lin2srgb_lut[1025];
static int lut_ready = 0;
/* color/srgb.go:57-75,86-93 */
static void lut_init(void) {
    if (lut_ready) return;
    for (int i = 0; i < 1024; i++) {
        double v = (double)i / 1024;
        lin2srgb_lut[i] = v <= 0.0031308 ? v * 12.92 : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
    }
    lin2srgb_lut[1024] = lin2srgb_lut[1023];
    lut_ready = 1;
}
/* color/srgb.go:15-31 */
static float from_linear_srgb(float v) {
    lut_init();
    if (v <= 0) return 0;
    if (v == 1) return 1;
    float i = v * 1024;
    int64_t ifl = gf2i(i) & 1023;
    float v0 = (float)lin2srgb_lut[ifl], v1 = (float)lin2srgb_lut[ifl + 1];
    i -= (float)ifl;
    return v0 * (1.0f - i) + v1 * i;
}
static float from_srgb_linear(float v) { /* color/srgb.go:34-50 with its LUT */
    if (v <= 0) return 0;
    if (v >= 1) return 1;
    float i = v * 1024;
    int64_t ifl = gf2i(i) & 1023;
    double l0 = (double)ifl / 1024, l1 = (double)(ifl + 1) / 1024;
    double t0 = l0 <= 0.04045 ? l0 / 12.92 : pow((l0 + 0.055) / 1.055, 2.4);
    double t1 = ifl + 1 < 1024 ? (l1 <= 0.04045 ? l1 / 12.92 : pow((l1 + 0.055) / 1.055, 2.4)) : t0;
    if (ifl + 1 == 1024) t1 = t0;
    float v0 = (float)t0, v1 = (float)t1;
    i -= (float)ifl;
    return v0 * (1.0f - i) + v1 * i;
}
/* shader/gamma.go:13-18 */
static uint8_t gamma_u8(uint8_t c) { return gf2u8(from_linear_srgb((float)c / 0xff) * 0xff + 0.5f); }

/* ----------------------------------------------------------------- resize */
/* imageutil.Resize (internal/imageutil/resize.go:16-66), the MSAA
 * antialiasing pass (render/raster.go:361-378): a separable bilinear filter
 * with 8-bit fixed-point weights; the first pass filters the rows of the
 * input into a transposed temporary, the second filters the rows of that. */
static float resize_linear(float in) { /* resize.go:68-74 */
    in = (float)fabs((double)in);
    if (in <= 1) return 1 - in;
    return 0;
}
static int resize_filter_len(float scale) { /* resize.go:149 with filterLength 2 */
    return 2 * (int)gmax2((float)ceil((double)scale), 1.0f);
}
/* createWeights8, resize.go:148-164 */
static void resize_weights(int dy, int fl, float scale, int16_t* coeffs, int* start) {
    const float ff = gmin2(1.0f / scale, 1.0f);
    for (int y = 0; y < dy; y++) {
        float interp = scale * ((float)y + 0.5f) - 0.5f;
        start[y] = (int)interp - fl / 2 + 1;
        interp -= (float)start[y];
        for (int i = 0; i < fl; i++) {
            float in = (interp - (float)i) * ff;
            coeffs[y * fl + i] = (int16_t)(int32_t)(resize_linear(in) * 256);
        }
    }
}
/* resizeRGBA, resize.go:76-118: `in` has in_rows rows of in_w RGBA pixels;
 * out[y][x] (out_rows x out_w, row-major) = filter of input row x at y. */
static void resize_rgba(const uint8_t* in, int in_w, uint8_t* out, int out_w, int out_rows, const int16_t* coeffs,
                        const int* start, int fl) {
    const int maxX = in_w - 1;
    for (int x = 0; x < out_w; x++) {
        const uint8_t* row = in + (size_t)x * in_w * 4;
        for (int y = 0; y < out_rows; y++) {
            int32_t rgba[4] = {0, 0, 0, 0}, sum = 0;
            for (int i = 0; i < fl; i++) {
                const int16_t coeff = coeffs[y * fl + i];
                if (coeff == 0) continue;
                int xi = start[y] + i;
                if ((unsigned)xi < (unsigned)maxX) xi *= 4;
                else if (xi >= maxX) xi = 4 * maxX;
                else xi = 0;
                for (int k = 0; k < 4; k++) rgba[k] += (int32_t)coeff * (int32_t)row[xi + k];
                sum += coeff;
            }
            for (int k = 0; k < 4; k++) {
                int v = rgba[k] / sum; /* Go's integer division truncates toward zero, as C's */
                out[((size_t)y * out_w + x) * 4 + k] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
            }
        }
    }
}
/* calcFactors (resize.go:121-138) and the size Resize then makes
 * (resize.go:17-23): a zero width or height follows the other's scale. */
static void resize_plan(int width, int height, int iw, int ih, float* sx, float* sy, int* ow, int* oh) {
    const float oldw = (float)iw, oldh = (float)ih;
    if (width == 0) {
        if (height == 0) {
            *sx = *sy = 1.0f;
        } else {
            *sy = oldh / (float)height;
            *sx = *sy;
        }
    } else {
        *sx = oldw / (float)width;
        *sy = height == 0 ? *sx : oldh / (float)height;
    }
    if (width == 0) width = (int)(0.7f + oldw / *sx);
    if (height == 0) height = (int)(0.7f + oldh / *sy);
    *ow = width, *oh = height;
}
/* Resize(width, height, img) with img (iw x ih RGBA8, top row first) into
 * out (the size resize_plan gives); returns 0, or -1 on allocation failure.
 * The identity when the sizes match. */
static int resize_image(const uint8_t* img, int iw, int ih, uint8_t* out, int width, int height) {
    float sx, sy;
    resize_plan(width, height, iw, ih, &sx, &sy, &width, &height);
    if ((iw == width && ih == height) || iw <= 0 || ih <= 0) {
        memcpy(out, img, (size_t)iw * ih * 4);
        return 0;
    }
    const int flx = resize_filter_len(sx), fly = resize_filter_len(sy);
    uint8_t* temp = malloc((size_t)ih * width * 4); /* image.Rect(0, 0, ih, width): width rows of ih pixels */
    int16_t* cx = malloc(sizeof(int16_t) * (size_t)width * flx);
    int16_t* cy = malloc(sizeof(int16_t) * (size_t)height * fly);
    int* ox = malloc(sizeof(int) * (size_t)width);
    int* oy = malloc(sizeof(int) * (size_t)height);
    int rc = -1;
    if (temp && cx && cy && ox && oy) {
        resize_weights(width, flx, sx, cx, ox);
        resize_rgba(img, iw, temp, ih, width, cx, ox, flx);
        resize_weights(height, fly, sy, cy, oy);
        resize_rgba(temp, ih, out, width, height, cy, oy, fly);
        rc = 0;
    }
    free(temp);
    free(cx);
    free(cy);
    free(ox);
    free(oy);
    return rc;
}

/* --------------------------------------------------------------- textures */
typedef struct {
    int w, h;
    uint8_t* px; /* [h][w][4] */
} img_t;
typedef struct {
    int use_mipmap, size, n;
    img_t lv[32];
} tex_t;

/* math.Log2 as Go computes it (Frexp, then Log of the fraction) */
static double go_log2(double x) {
    if (!(x < INFINITY)) return x; /* Log2(+Inf) = +Inf, Log2(NaN) = NaN */
    int e;
    double f = frexp(x, &e);
    if (f == 0.5) return (double)(e - 1);
    return log(f) * (1 / M_LN2) + (double)e;
}

/* buffer.NewTexture (buffer/texture.go:35-68): level i is
 * Resize(dx / 2^i, dy / 2^i) of the image, for int(Log2(max(dx, dy))) + 1
 * levels; a 1x1 image has one. */
static int tex_build(const prh_texture* t, tex_t* out) {
    memset(out, 0, sizeof *out);
    out->use_mipmap = t->use_mipmap != 0;
    out->size = t->width;
    const int dx = t->width, dy = t->height;
    out->lv[0].w = dx, out->lv[0].h = dy;
    out->lv[0].px = malloc((size_t)dx * dy * 4);
    if (!out->lv[0].px) return -1;
    memcpy(out->lv[0].px, t->rgba, (size_t)dx * dy * 4);
    out->n = 1;
    if (dx == 1 && dy == 1) return 0;
    int L = (int)(float)go_log2((double)gmax2((float)dx, (float)dy)) + 1;
    if (L > 32) L = 32;
    for (int i = 1; i < L; i++) {
        const int d = (int)gpow(2.0f, (float)i);
        float sx, sy;
        int w, h;
        resize_plan(dx / d, dy / d, dx, dy, &sx, &sy, &w, &h);
        out->lv[i].w = w, out->lv[i].h = h;
        out->lv[i].px = malloc((size_t)w * h * 4 + 4);
        if (!out->lv[i].px || resize_image(t->rgba, dx, dy, out->lv[i].px, dx / d, dy / d)) return -1;
        out->n = i + 1;
    }
    return 0;
}
static void tex_free(tex_t* t) {
    for (int i = 0; i < t->n; i++) free(t->lv[i].px);
    t->n = 0;
}
/* math.Modf as Go computes it: Modf(+-Inf) = +-Inf, NaN (C's modf gives 0) */
static double go_modf(double x, double* ip) {
    if (isinf(x)) {
        *ip = x;
        return NAN;
    }
    return modf(x, ip);
}
/* image.RGBA.RGBAAt: transparent black outside the image (indices are Go
 * ints: int(Floor(NaN)) is the most negative one, outside) */
static void rgba_at(const img_t* im, int64_t x, int64_t y, uint8_t o[4]) {
    if (x < 0 || y < 0 || x >= im->w || y >= im->h) {
        memset(o, 0, 4);
        return;
    }
    memcpy(o, im->px + 4 * ((size_t)y * im->w + x), 4);
}
/* math.LerpC (math/interpolate.go:55-62) */
static void lerp_c(const uint8_t a[4], const uint8_t b[4], float t, uint8_t o[4]) {
    for (int k = 0; k < 4; k++) o[k] = gf2u8((float)a[k] + t * ((float)b[k] - (float)a[k]));
}
/* queryBilinear, buffer/texture.go:162-195 */
static void tex_bilinear(const tex_t* t, int lod, float u, float v, uint8_t o[4]) {
    const img_t* b = &t->lv[lod];
    const int dx = b->w, dy = b->h;
    if (dx == 1 && dy == 1) {
        rgba_at(b, 0, 0, o);
        return;
    }
    const float x = u * ((float)dx - 1), y = v * ((float)dy - 1);
    const float x0 = (float)floor((double)x), y0 = (float)floor((double)y);
    const int64_t i = gf2i(x0), j = gf2i(y0);
    uint8_t p1[4], p2[4], p3[4], p4[4], q1[4], q2[4];
    rgba_at(b, i, j, p1);
    if (i < dx - 1) rgba_at(b, i + 1, j, p2);
    else rgba_at(b, i, j, p2);
    lerp_c(p1, p2, x - x0, q1);
    if (j < dy - 1) rgba_at(b, i, j + 1, p3);
    else rgba_at(b, i, j, p3);
    if (i < dx - 1 && j < dy - 1) rgba_at(b, i + 1, j + 1, p4);
    else rgba_at(b, i, j, p4);
    lerp_c(p3, p4, x - x0, q2);
    lerp_c(q1, q2, y - y0, o);
}
/* Texture.Query, buffer/texture.go:83-154 */
static void tex_query(const tex_t* t, float lod, float u, float v, uint8_t o[4]) {
    double iu, iv;
    float fu = (float)go_modf((double)u, &iu), fv = (float)go_modf((double)v, &iv);
    if ((float)iu != 0 && fu == 0) fu = 1;
    if (fu < 0) fu = 1 - fu;
    if ((float)iv != 0 && fv == 0) fv = 1;
    if (fv < 0) fv = 1 - fv;
    if (!t->use_mipmap) { /* queryL0, texture.go:130-141 */
        const img_t* b = &t->lv[0];
        if (b->w == 1 && b->h == 1) {
            rgba_at(b, 0, 0, o);
            return;
        }
        const int64_t x = gf2i((float)floor((double)(fu * ((float)b->w - 1))));
        const int64_t y = gf2i((float)floor((double)(fv * ((float)b->h - 1))));
        rgba_at(b, x, y, o);
        return;
    }
    if (lod < 0) lod = 0;
    else if (lod >= (float)t->n) lod = (float)(t->n - 1);
    if (lod <= 1) {
        tex_bilinear(t, 0, fu, fv, o);
        return;
    }
    lod -= 1;
    const int h = (int)gf2i((float)floor((double)lod)), l = h + 1;
    if (l >= t->n) {
        tex_bilinear(t, h, fu, fv, o);
        return;
    }
    const float p = lod - (float)h;
    if (approx_eq(p, 0, 1e-7f)) {
        tex_bilinear(t, h, fu, fv, o);
        return;
    }
    uint8_t a[4], b[4];
    tex_bilinear(t, h, fu, fv, a);
    tex_bilinear(t, l, fu, fv, b);
    lerp_c(a, b, p, o);
}
/* FragmentShader's texture lookup (shader/blinn_cpu.go:28-35) */
static void tex_fragment(const tex_t* t, float du, float dv, float u, float v, uint8_t o[4]) {
    float lod = 0;
    if (t->use_mipmap) {
        float siz = (float)t->size * gsqrt(gmax2(du, dv));
        if (siz < 1) siz = 1;
        lod = (float)go_log2((double)siz);
    }
    tex_query(t, lod, u, 1 - v, o);
}
int oracle_texture_query(const prh_texture* tx, float lod, float u, float v, uint8_t out[4]) {
    tex_t t;
    if (tex_build(tx, &t)) {
        tex_free(&t);
        return PRH_E_NOMEM;
    }
    tex_query(&t, lod, u, v, out);
    tex_free(&t);
    return PRH_OK;
}
int oracle_texture_level(const prh_texture* tx, int level, int* w, int* h, uint8_t* out, size_t bytes) {
    tex_t t;
    if (tex_build(tx, &t)) {
        tex_free(&t);
        return PRH_E_NOMEM;
    }
    int rc = PRH_E_INVALID;
    if (level >= 0 && level < t.n) {
        *w = t.lv[level].w, *h = t.lv[level].h;
        if (out && bytes >= (size_t)*w * *h * 4) memcpy(out, t.lv[level].px, (size_t)*w * *h * 4);
        rc = t.n;
    }
    tex_free(&t);
    return rc;
}
int oracle_resize(const uint8_t* img, int iw, int ih, uint8_t* out, int width, int height) {
    return resize_image(img, iw, ih, out, width, height) ? PRH_E_NOMEM : PRH_OK;
}

/* ------------------------------------------------------------ frame state */
typedef struct {
    int ok;
    int x, y;
    float depth;
    v4 nor, facenor, wpos;
    uint8_t col[4];
    float u, v, du, dv;
    int64_t matid;
    uint32_t order; /* oracle bookkeeping: triangle*8+sub of the stored fragment */
} frag_t;

typedef struct {
    v4 pos, nor;
    uint8_t col[4];
    float uv[2];
} vert_t;

typedef struct {
    const prh_scene_desc* s;
    const prh_camera* cam;
    const prh_options* o;
    int W, H;         /* framebuffer (W*MSAA, H*MSAA) */
    int perspect;
    frag_t* f;        /* storage rows (H-1-y) as FragmentBuffer */
    /* shadow */
    int SW, SH;
    m4* lview;
    m4* lproj;
    int* lactive;
    float* smaps;
    /* deferred uniforms */
    m4 view, proj, vp, viewinv, projinv, vpinv, s2w;
    float ao_dir[16];
    float* linear;
    uint8_t* shadow_count;
    float* ao_total;
    double shade_err64;
    tex_t* tex; /* the scene's textures with their mip levels */
    int n_tex;
    /* throughput mode (oracle_set_workers > 1, the CPU baseline only):
     * triangles drawn concurrently, each pixel's test and set under a spin
     * lock as the reference's FragmentBuffer does (buffer/buffer.go:209-280) */
    int workers;
    uint8_t* locks;
} ctx_t;

/* 0: the parity order (sequential raster, deferred pass on OpenMP's threads);
 * 1: everything on one thread (the reference's Workers(1), Mode A);
 * n > 1: n workers for the raster passes and the deferred pass (Mode B; tie
 * order between concurrent triangles is then arbitrary, as in the
 * reference with several workers). */
static int g_workers = 0;
void oracle_set_workers(int n) { g_workers = n < 0 ? 0 : n; }

static inline frag_t* fat(ctx_t* c, int x, int y) { return &c->f[(size_t)(c->H - 1 - y) * c->W + x]; }
static inline int in_buf(const ctx_t* c, int64_t x, int64_t y) { return x >= 0 && x < c->W && y >= 0 && y < c->H; }

static inline void px_lock(ctx_t* c, size_t i) {
    while (__atomic_test_and_set(&c->locks[i], __ATOMIC_ACQUIRE)) {
    }
}
static inline void px_unlock(ctx_t* c, size_t i) { __atomic_clear(&c->locks[i], __ATOMIC_RELEASE); }

/* buffer/buffer.go:269-280 */
static inline int depth_test(ctx_t* c, int x, int y, float z) {
    frag_t* f = fat(c, x, y);
    if (!c->locks) return !f->ok || z > f->depth;
    size_t i = (size_t)(c->H - 1 - y) * c->W + x;
    px_lock(c, i);
    int r = !f->ok || z > f->depth;
    px_unlock(c, i);
    return r;
}
/* buffer/buffer.go:221-266 (colour/depth images are produced later) */
static inline void fset(ctx_t* c, int x, int y, const frag_t* nf) {
    frag_t* f = fat(c, x, y);
    size_t i = (size_t)(c->H - 1 - y) * c->W + x;
    if (c->locks) px_lock(c, i);
    if (!(f->ok && nf->depth <= f->depth)) *f = *nf;
    if (c->locks) px_unlock(c, i);
}

/* render/raster.go:453-460 */
static inline v4 interp_wpos(const float bc[3], v4 m1, v4 m2, v4 m3) {
    return V4(bc[0] * m1.x + bc[1] * m2.x + bc[2] * m3.x, bc[0] * m1.y + bc[1] * m2.y + bc[2] * m3.y,
              bc[0] * m1.z + bc[1] * m2.z + bc[2] * m3.z, 1);
}

/* render/raster.go:462-572 */
static void draw_clipped(ctx_t* c, const m4* vpinv, const m4* pinv, const m4* vinv, const vert_t* t1,
                         const vert_t* t2, const vert_t* t3, const float recipw[3], int64_t matid, uint32_t order) {
    v4 m1 = apply4(apply4(apply4(t1->pos, vpinv), pinv), vinv);
    v4 m2 = apply4(apply4(apply4(t2->pos, vpinv), pinv), vinv);
    v4 m3 = apply4(apply4(apply4(t3->pos, vpinv), pinv), vinv);
    v3 ps[3] = {V3(t1->pos.x, t1->pos.y, t1->pos.z), V3(t2->pos.x, t2->pos.y, t2->pos.z),
                V3(t3->pos.x, t3->pos.y, t3->pos.z)};
    aabb_t bb = aabb_new(3, ps);
    int64_t xmin = gf2i(ground(bb.mn.x) - 1), xmax = gf2i(ground(bb.mx.x) + 1);
    int64_t ymin = gf2i(ground(bb.mn.y) - 1), ymax = gf2i(ground(bb.mx.y) + 1);
    v4 fN = unit4(cross4(sub4(m2, m1), sub4(m3, m1)));
    if (xmin < 0) xmin = 0;
    if (ymin < 0) ymin = 0;
    if (xmax > c->W - 1) xmax = c->W - 1;
    if (ymax > c->H - 1) ymax = c->H - 1;
    for (int64_t x = xmin; x <= xmax; x++) {
        for (int64_t y = ymin; y <= ymax; y++) {
            float bc[3];
            barycoord((float)x + 0.5f, (float)y + 0.5f, t1->pos.x, t1->pos.y, t2->pos.x, t2->pos.y, t3->pos.x,
                      t3->pos.y, bc);
            if (bc[0] < -1e-7f || bc[1] < -1e-7f || bc[2] < -1e-7f) continue;
            float z = bc[0] * t1->pos.z + bc[1] * t2->pos.z + bc[2] * t3->pos.z;
            if (!depth_test(c, (int)x, (int)y, z)) continue;
            float wc1 = recipw[0] * bc[0], wc2 = recipw[1] * bc[1], wc3 = recipw[2] * bc[2];
            float norm = 1.0f;
            if (c->perspect) norm = 1 / (wc1 + wc2 + wc3);
            frag_t nf;
            nf.ok = 1;
            nf.x = (int)x;
            nf.y = (int)y;
            nf.depth = z;
            nf.nor = unit4(V4(bc[0] * t1->nor.x + bc[1] * t2->nor.x + bc[2] * t3->nor.x,
                              bc[0] * t1->nor.y + bc[1] * t2->nor.y + bc[2] * t3->nor.y,
                              bc[0] * t1->nor.z + bc[1] * t2->nor.z + bc[2] * t3->nor.z, 0));
            nf.facenor = fN;
            nf.wpos = interp_wpos(bc, m1, m2, m3);
            /* UV and its screen derivatives (raster.go:511-535); the latter's
             * normalisation is unconditional, unlike the UV's */
            const float px = (float)x + 0.5f, py = (float)y + 0.5f;
            nf.u = (wc1 * t1->uv[0] + wc2 * t2->uv[0] + wc3 * t3->uv[0]) * norm;
            nf.v = (wc1 * t1->uv[1] + wc2 * t2->uv[1] + wc3 * t3->uv[1]) * norm;
            nf.du = nf.dv = 0;
            if (matid >= 0) {
                float bx[3], by[3];
                barycoord(px + 1, py, t1->pos.x, t1->pos.y, t2->pos.x, t2->pos.y, t3->pos.x, t3->pos.y, bx);
                float w1x = recipw[0] * bx[0], w2x = recipw[1] * bx[1], w3x = recipw[2] * bx[2];
                float nx = 1 / (w1x + w2x + w3x);
                barycoord(px, py + 1, t1->pos.x, t1->pos.y, t2->pos.x, t2->pos.y, t3->pos.x, t3->pos.y, by);
                float w1y = recipw[0] * by[0], w2y = recipw[1] * by[1], w3y = recipw[2] * by[2];
                float ny = 1 / (w1y + w2y + w3y);
                float uU = (w1x * t1->uv[0] + w2x * t2->uv[0] + w3x * t3->uv[0]) * nx;
                float uX = (w1x * t1->uv[1] + w2x * t2->uv[1] + w3x * t3->uv[1]) * nx;
                float vV = (w1y * t1->uv[0] + w2y * t2->uv[0] + w3y * t3->uv[0]) * ny;
                float vY = (w1y * t1->uv[1] + w2y * t2->uv[1] + w3y * t3->uv[1]) * ny;
                nf.du = (uU - nf.u) * (uU - nf.u) + (uX - nf.v) * (uX - nf.v);
                nf.dv = (vV - nf.u) * (vV - nf.u) + (vY - nf.v) * (vY - nf.v);
            }
            for (int k = 0; k < 4; k++)
                nf.col[k] = gf2u8(gclamp(
                    (wc1 * (float)t1->col[k] + wc2 * (float)t2->col[k] + wc3 * (float)t3->col[k]) * norm, 0, 0xff));
            nf.matid = matid;
            nf.order = order;
            fset(c, (int)x, (int)y, &nf);
        }
    }
}

/* render/clipping.go:14-65 */
typedef struct { v4 pos, nor; } plane_t;
static inline int in_front(const plane_t* p, v4 v) { return dot4(sub4(v, p->pos), p->nor) > 0; }
static inline v4 isect(const plane_t* p, v4 v0, v4 v1) {
    v4 u = sub4(v1, v0), w = sub4(v0, p->pos);
    float d = dot4(p->nor, u);
    float n = -dot4(p->nor, w);
    float s = n / d;
    return add4(v0, scale4(u, s, s, s, s));
}
static int sutherland_hodgman(const v4* pts, int n, float w, float h, v4* out) {
    plane_t planes[6] = {{V4(w, 0, 0, 1), V4(-1, 0, 0, 1)}, {V4(0, 0, 0, 1), V4(1, 0, 0, 1)},
                         {V4(0, h, 0, 1), V4(0, -1, 0, 1)}, {V4(0, 0, 0, 1), V4(0, 1, 0, 1)},
                         {V4(0, 0, 1, 1), V4(0, 0, -1, 1)}, {V4(0, 0, -1, 1), V4(0, 0, 1, 1)}};
    v4 a[16], b[16];
    int na = n;
    for (int i = 0; i < n; i++) a[i] = pts[i];
    for (int pi = 0; pi < 6; pi++) {
        if (na == 0) return 0;
        int nb = 0;
        v4 s = a[na - 1];
        for (int i = 0; i < na; i++) {
            v4 e = a[i];
            if (in_front(&planes[pi], e)) {
                if (!in_front(&planes[pi], s)) b[nb++] = isect(&planes[pi], s, e);
                b[nb++] = e;
            } else if (in_front(&planes[pi], s)) {
                b[nb++] = isect(&planes[pi], s, e);
            }
            s = e;
        }
        for (int i = 0; i < nb; i++) a[i] = b[i];
        na = nb;
    }
    for (int i = 0; i < na; i++) out[i] = a[i];
    return na;
}
/* render/clipping.go:67-157 */
static int clip_triangle(const vert_t* v1, const vert_t* v2, const vert_t* v3, float w, float h, vert_t (*res)[3]) {
    v4 pts[3] = {v1->pos, v2->pos, v3->pos}, clips[16];
    int n = sutherland_hodgman(pts, 3, w, h, clips);
    int nt = 0;
    for (int i = 2; i < n; i++) {
        const v4* cp[3] = {&clips[0], &clips[i - 1], &clips[i]};
        for (int k = 0; k < 3; k++) {
            float b[3];
            barycoord(cp[k]->x, cp[k]->y, v1->pos.x, v1->pos.y, v2->pos.x, v2->pos.y, v3->pos.x, v3->pos.y, b);
            vert_t* t = &res[nt][k];
            t->pos = V4(b[0] * v1->pos.x + b[1] * v2->pos.x + b[2] * v3->pos.x,
                        b[0] * v1->pos.y + b[1] * v2->pos.y + b[2] * v3->pos.y,
                        b[0] * v1->pos.z + b[1] * v2->pos.z + b[2] * v3->pos.z, 1);
            t->uv[0] = b[0] * v1->uv[0] + b[1] * v2->uv[0] + b[2] * v3->uv[0];
            t->uv[1] = b[0] * v1->uv[1] + b[1] * v2->uv[1] + b[2] * v3->uv[1];
            t->nor = V4(b[0] * v1->nor.x + b[1] * v2->nor.x + b[2] * v3->nor.x,
                        b[0] * v1->nor.y + b[1] * v2->nor.y + b[2] * v3->nor.y,
                        b[0] * v1->nor.z + b[1] * v2->nor.z + b[2] * v3->nor.z, 0);
            for (int q = 0; q < 4; q++)
                t->col[q] = gf2u8(gclamp(
                    b[0] * (float)v1->col[q] + b[1] * (float)v2->col[q] + b[2] * (float)v3->col[q], 0, 0xff));
        }
        nt++;
    }
    return nt;
}

/* render/cull.go:26-28 */
static inline int cull_back_face(v4 v1, v4 v2, v4 v3) { return cross4(sub4(v2, v1), sub4(v3, v1)).z < 0; }

static void load_vert(const prh_scene_desc* s, int vi, vert_t* v) {
    v->pos = V4(s->vpos[4 * vi], s->vpos[4 * vi + 1], s->vpos[4 * vi + 2], s->vpos[4 * vi + 3]);
    v->nor = V4(s->vnor[4 * vi], s->vnor[4 * vi + 1], s->vnor[4 * vi + 2], s->vnor[4 * vi + 3]);
    memcpy(v->col, s->vcol + 4 * vi, 4);
    v->uv[0] = s->vuv[2 * vi];
    v->uv[1] = s->vuv[2 * vi + 1];
}
static m4 obj_model(const prh_scene_desc* s, int o) {
    m4 m;
    memcpy(m.a, s->obj_model + 16 * o, sizeof m.a);
    return m;
}

/* The first global triangle of every object (n_objects + 1 entries). */
static int64_t* obj_tri_starts(const prh_scene_desc* s) {
    int64_t* st = malloc(sizeof(int64_t) * ((size_t)s->n_objects + 1));
    if (!st) return NULL;
    st[0] = 0;
    for (int o = 0; o < s->n_objects; o++) {
        int mesh = s->obj_mesh[o];
        st[o + 1] = st[o] + (s->mesh_tri_off[mesh + 1] - s->mesh_tri_off[mesh]);
    }
    return st;
}
static int obj_of(const int64_t* st, int n_objects, int64_t g) {
    int lo = 0, hi = n_objects - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (st[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static void forward_tri(ctx_t* c, const m4* vp, const m4* vpinv, const m4* pinv, const m4* vinv, const m4* trans,
                        const m4* normal, int o, int v0, int t, int64_t gtri);

/* render/raster.go:226-273 and draw, render/raster.go:380-444.  In
 * throughput mode (workers > 1) the triangles are drawn concurrently. */
static void forward_pass(ctx_t* c) {
    const prh_scene_desc* s = c->s;
    m4 view = c->view, proj = c->proj, vp = c->vp;
    m4 vinv = inv(&view), pinv = inv(&proj), vpinv = inv(&vp);
    int64_t* st = c->workers > 1 ? obj_tri_starts(s) : NULL;
    if (st) {
        m4 pv = mulm(&proj, &view);
#pragma omp parallel for schedule(dynamic, 256) num_threads(c->workers)
        for (int64_t g = 0; g < st[s->n_objects]; g++) {
            int o = obj_of(st, s->n_objects, g);
            m4 model = obj_model(s, o);
            m4 mi = inv(&model);
            m4 normal = transpose(&mi);
            m4 trans = mulm(&pv, &model);
            int mesh = s->obj_mesh[o];
            int t = s->mesh_tri_off[mesh] + (int)(g - st[o]);
            forward_tri(c, &vp, &vpinv, &pinv, &vinv, &trans, &normal, o, s->mesh_vtx_off[mesh], t, g);
        }
        free(st);
        return;
    }
    int64_t gtri = 0;
    for (int o = 0; o < s->n_objects; o++) {
        int mesh = s->obj_mesh[o];
        m4 model = obj_model(s, o);
        m4 mi = inv(&model);
        m4 normal = transpose(&mi);
        m4 pv = mulm(&proj, &view);
        m4 trans = mulm(&pv, &model);
        int v0 = s->mesh_vtx_off[mesh];
        for (int t = s->mesh_tri_off[mesh]; t < s->mesh_tri_off[mesh + 1]; t++, gtri++)
            forward_tri(c, &vp, &vpinv, &pinv, &vinv, &trans, &normal, o, v0, t, gtri);
    }
}

/* draw (render/raster.go:380-444) of mesh triangle t of object o, global
 * triangle gtri */
static void forward_tri(ctx_t* c, const m4* vp, const m4* vpinv, const m4* pinv, const m4* vinv, const m4* trans,
                        const m4* normal, int o, int v0, int t, int64_t gtri) {
    const prh_scene_desc* s = c->s;
    int64_t mat = s->tri_mat[t];
    if (mat >= 0) mat += s->obj_mat_base[o];
    vert_t in[3], tv[3];
    for (int k = 0; k < 3; k++) load_vert(s, v0 + s->tri_v[3 * t + k], &in[k]);
    if (!tri_is_valid(in[0].pos, in[1].pos, in[2].pos)) return;
    for (int k = 0; k < 3; k++) {
        tv[k] = in[k];
        tv[k].pos = mulv(trans, in[k].pos);
        tv[k].nor = apply4(in[k].nor, normal);
    }
    float recipw[3] = {1, 1, 1};
    if (c->perspect)
        for (int k = 0; k < 3; k++) recipw[k] = -1 / tv[k].pos.w;
    for (int k = 0; k < 3; k++) tv[k].pos = pos4(apply4(tv[k].pos, vp));
    if (cull_back_face(tv[0].pos, tv[1].pos, tv[2].pos)) return;
    v3 p[3];
    for (int k = 0; k < 3; k++) p[k] = V3(tv[k].pos.x, tv[k].pos.y, tv[k].pos.z);
    int msaa = c->o->msaa > 0 ? c->o->msaa : 1;
    aabb_t vpbox = {V3(0, 0, -1), V3((float)(msaa * c->W), (float)(msaa * c->H), 1)};
    aabb_t tbox = aabb_new(3, p);
    if (!aabb_intersect(&vpbox, &tbox)) return;
    uint32_t key = (uint32_t)(gtri * 8);
    if (aabb_contains(&vpbox, 3, p)) {
        draw_clipped(c, vpinv, pinv, vinv, &tv[0], &tv[1], &tv[2], recipw, mat, key + 7);
        return;
    }
    vert_t sub[8][3];
    int ns = clip_triangle(&tv[0], &tv[1], &tv[2], (float)(msaa * c->W), (float)(msaa * c->H), sub);
    for (int i = 0; i < ns; i++)
        draw_clipped(c, vpinv, pinv, vinv, &sub[i][0], &sub[i][1], &sub[i][2], recipw, mat, key + i);
}

/* ------------------------------------------------------------------ shadows */
/* render/shadow.go:33-90.  A shadow-casting directional light gets the same
 * orthographic light camera as a point light (the reference leaves it nil). */
static void init_shadow_maps(ctx_t* c) {
    const prh_scene_desc* s = c->s;
    int L = s->n_lights;
    c->lview = calloc((size_t)L + 1, sizeof(m4));
    c->lproj = calloc((size_t)L + 1, sizeof(m4));
    c->lactive = calloc((size_t)L + 1, sizeof(int));
    c->smaps = calloc((size_t)L * c->SW * c->SH + 1, sizeof(float));
    v3 up = V3(0, 1, 0);
    v3 center = V3a(s->center);
    m4 cview = view_matrix(V3a(c->cam->position), V3a(c->cam->target), V3a(c->cam->up));
    m4 cproj = cam_proj(c->cam);
    m4 cvinv = inv(&cview), cpinv = inv(&cproj);
    for (int i = 0; i < L; i++) {
        if (!s->lights[i].cast_shadow) continue;
        v3 lp = V3a(s->lights[i].position);
        m4 lv = view_matrix(lp, center, up);
        m4 t0 = mulm(&lv, &cvinv);
        m4 tm = mulm(&t0, &cpinv);
        static const float corners[8][3] = {{1, 1, 1},  {1, 1, -1},  {1, -1, 1},  {-1, 1, 1},
                                            {-1, -1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};
        v3 cv[8];
        for (int k = 0; k < 8; k++) {
            v4 q = pos4(apply4(V4(corners[k][0], corners[k][1], corners[k][2], 1), &tm));
            cv[k] = V3(q.x, q.y, q.z);
        }
        aabb_t bb = aabb_new(8, cv);
        float le = bb.mn.x, ri = bb.mx.x, bo = bb.mn.y, to = bb.mx.y, ne = bb.mx.z, fa = bb.mn.z - 2;
        c->lview[i] = view_matrix(lp, center, up);
        c->lproj[i] = ortho_proj(le, ri, bo, to, ne, fa);
        c->lactive[i] = 1;
    }
}

/* render/shadow.go:152-228 */
static void draw_depth(ctx_t* c, int li, const m4* trans, const m4* vp, const vert_t* in) {
    v4 p[3];
    for (int k = 0; k < 3; k++) p[k] = pos4(apply4(mulv(trans, in[k].pos), vp));
    if (cull_back_face(p[0], p[1], p[2])) return;
    /* cullViewFrustum, render/cull.go:15-23 */
    int msaa = c->o->msaa > 0 ? c->o->msaa : 1;
    v3 vpp[3] = {V3((float)(c->SW * msaa), (float)(c->SH * msaa), 1), V3(0, 0, 0), V3(0, 0, -1)};
    aabb_t vpbox = aabb_new(3, vpp);
    v3 q[3];
    for (int k = 0; k < 3; k++) q[k] = V3(p[k].x, p[k].y, p[k].z);
    aabb_t tb = aabb_new(3, q);
    if (!aabb_intersect(&vpbox, &tb)) return;
    int64_t xmin = gf2i(ground(tb.mn.x) - 1), xmax = gf2i(ground(tb.mx.x) + 1);
    int64_t ymin = gf2i(ground(tb.mn.y) - 1), ymax = gf2i(ground(tb.mx.y) + 1);
    if (xmin < 0) xmin = 0;
    if (ymin < 0) ymin = 0;
    if (xmax > c->SW - 1) xmax = c->SW - 1;
    if (ymax > c->SH - 1) ymax = c->SH - 1;
    float* d = c->smaps + (size_t)li * c->SW * c->SH;
    for (int64_t x = xmin; x <= xmax; x++)
        for (int64_t y = ymin; y <= ymax; y++) {
            float bc[3];
            barycoord((float)x + 0.5f, (float)y + 0.5f, p[0].x, p[0].y, p[1].x, p[1].y, p[2].x, p[2].y, bc);
            if (bc[0] < -1e-7f || bc[1] < -1e-7f || bc[2] < -1e-7f) continue;
            float z = bc[0] * p[0].z + bc[1] * p[1].z + bc[2] * p[2].z;
            size_t idx = (size_t)(x + y * c->SW);
            if (!c->locks) {
                if (!(z <= d[idx])) d[idx] = z;
                continue;
            }
            /* throughput mode: the same keep-the-larger store as one atomic step */
            uint32_t* cell = (uint32_t*)&d[idx];
            uint32_t old = __atomic_load_n(cell, __ATOMIC_RELAXED), nz;
            memcpy(&nz, &z, 4);
            for (;;) {
                float of;
                memcpy(&of, &old, 4);
                if (z <= of) break;
                if (__atomic_compare_exchange_n(cell, &old, nz, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) break;
            }
        }
}
/* render/shadow.go:92-150 */
static void pass_shadows(ctx_t* c, int li) {
    const prh_scene_desc* s = c->s;
    if (!s->lights[li].cast_shadow) return;
    m4 vp = viewport_matrix((float)c->SW, (float)c->SH);
    m4 pv = mulm(&c->lproj[li], &c->lview[li]);
    int64_t* st = c->workers > 1 ? obj_tri_starts(s) : NULL;
    if (st) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(c->workers)
        for (int64_t g = 0; g < st[s->n_objects]; g++) {
            int o = obj_of(st, s->n_objects, g);
            int mesh = s->obj_mesh[o];
            m4 model = obj_model(s, o);
            m4 trans = mulm(&pv, &model);
            int t = s->mesh_tri_off[mesh] + (int)(g - st[o]), v0 = s->mesh_vtx_off[mesh];
            vert_t in[3];
            for (int k = 0; k < 3; k++) load_vert(s, v0 + s->tri_v[3 * t + k], &in[k]);
            if (tri_is_valid(in[0].pos, in[1].pos, in[2].pos)) draw_depth(c, li, &trans, &vp, in);
        }
        free(st);
        return;
    }
    for (int o = 0; o < s->n_objects; o++) {
        int mesh = s->obj_mesh[o];
        m4 model = obj_model(s, o);
        m4 trans = mulm(&pv, &model);
        int v0 = s->mesh_vtx_off[mesh];
        for (int t = s->mesh_tri_off[mesh]; t < s->mesh_tri_off[mesh + 1]; t++) {
            vert_t in[3];
            for (int k = 0; k < 3; k++) load_vert(s, v0 + s->tri_v[3 * t + k], &in[k]);
            if (!tri_is_valid(in[0].pos, in[1].pos, in[2].pos)) continue;
            draw_depth(c, li, &trans, &vp, in);
        }
    }
}
/* render/shadow.go:243-282: the lookup of one fragment (x, y, depth) in one
 * shadow map: screen -> world -> light view -> light projection -> light
 * viewport, Pos(); texel int(x) + int(y)*SW with Go's truncating
 * conversions and wrapping int arithmetic; occluded iff 0 < idx < len and
 * z < map[idx] - 0.03. */
static int shadow_lookup(const m4* s2w, const m4* lview, const m4* lproj, int SW, int SH, const float* map, float x,
                         float y, float depth) {
    m4 vps = viewport_matrix((float)SW, (float)SH);
    v4 sc = V4(x, y, depth, 1);
    sc = apply4(sc, s2w);
    sc = apply4(sc, lview);
    sc = apply4(sc, lproj);
    sc = pos4(apply4(sc, &vps));
    int64_t lx = gf2i(sc.x), ly = gf2i(sc.y);
    int64_t idx = (int64_t)((uint64_t)lx + (uint64_t)ly * (uint64_t)(int64_t)SW);
    int64_t len = (int64_t)SW * SH;
    int shadow = 0;
    if (idx > 0 && idx < len) {
        float sz = map[idx];
        if (sc.z < sz - 0.03f) shadow++;
    }
    return shadow > 0;
}
/* render/shadow.go:230-283: true when the fragment is occluded (the name is
 * inverted in the reference; non-casting lights report true). */
static int shading_visibility(ctx_t* c, int li, const frag_t* info) {
    if (!c->s->lights[li].cast_shadow) return 1;
    return shadow_lookup(&c->s2w, &c->lview[li], &c->lproj[li], c->SW, c->SH,
                         c->smaps + (size_t)li * c->SW * c->SH, (float)info->x, (float)info->y, info->depth);
}

/* -------------------------------------------------------------- shading */
static const prh_material* mat_at(const prh_scene_desc* s, int64_t id) { /* render/material_table.go */
    if (id < 0 || id >= s->n_materials) return NULL;
    return &s->materials[id];
}

/* shader/blinn_cpu.go:24-106; lin receives the pre-Round sums.  Returns 0 on
 * the no-lights early return (texture colour). */
static int fragment_shader(const prh_material* m, const uint8_t* col, v4 nor, v4 facenor, v4 x, v3 cam, int nl,
                           const prh_light* ls, int na, const float* amb, uint8_t out[4], float lin[3],
                           double lin64[3]) {
    if (nl == 0) {
        memcpy(out, col, 4);
        return 0;
    }
    float LaR = 0, LaG = 0, LaB = 0;
    for (int e = 0; e < na; e++) {
        LaR += amb[e] * (float)col[0];
        LaG += amb[e] * (float)col[1];
        LaB += amb[e] * (float)col[2];
    }
    float LdR = 0, LdG = 0, LdB = 0, LsR = 0, LsG = 0, LsB = 0;
    v4 n = m->flat_shading ? facenor : nor;
    double dLa[3] = {0, 0, 0}, dLd[3] = {0, 0, 0}, dLs[3] = {0, 0, 0};
    for (int e = 0; e < na && lin64; e++)
        for (int k = 0; k < 3; k++) dLa[k] += (double)amb[e] * col[k];
    for (int li = 0; li < nl; li++) {
        const prh_light* l = &ls[li];
        v4 L;
        float I;
        if (l->kind == PRH_LIGHT_POINT) {
            v4 ldir = sub4(V4(l->position[0], l->position[1], l->position[2], 1), x);
            L = unit4(ldir);
            I = l->intensity / len4(ldir);
        } else {
            L = scale4(V4(l->direction[0], l->direction[1], l->direction[2], 0), -1, -1, -1, 1);
            I = l->intensity;
        }
        v4 V = unit4(sub4(V4(cam.x, cam.y, cam.z, 1), x));
        v4 H = unit4(add4(L, V));
        float Ld = gclamp(dot4(n, L), 0, 1);
        float Ls = gpow(gclamp(dot4(n, H), 0, 1), m->shininess);
        LdR += Ld * (float)col[0] * I;
        LdG += Ld * (float)col[1] * I;
        LdB += Ld * (float)col[2] * I;
        LsR += Ls * (float)l->color[0] * I;
        LsG += Ls * (float)l->color[1] * I;
        LsB += Ls * (float)l->color[2] * I;
        if (lin64) { /* the same shading, every operation in double */
            double nx = n.x, ny = n.y, nz = n.z, Lx, Ly, Lz, dI;
            if (l->kind == PRH_LIGHT_POINT) {
                double dx = (double)l->position[0] - x.x, dy = (double)l->position[1] - x.y,
                       dz = (double)l->position[2] - x.z;
                double dl = sqrt(dx * dx + dy * dy + dz * dz);
                Lx = dx / dl, Ly = dy / dl, Lz = dz / dl;
                dI = l->intensity / dl;
            } else {
                Lx = -(double)l->direction[0], Ly = -(double)l->direction[1], Lz = -(double)l->direction[2];
                dI = l->intensity;
            }
            double vx = (double)cam.x - x.x, vy = (double)cam.y - x.y, vz = (double)cam.z - x.z;
            double vl = sqrt(vx * vx + vy * vy + vz * vz);
            vx /= vl, vy /= vl, vz /= vl;
            double hx = Lx + vx, hy = Ly + vy, hz = Lz + vz, hl = sqrt(hx * hx + hy * hy + hz * hz);
            hx /= hl, hy /= hl, hz /= hl;
            double dld = nx * Lx + ny * Ly + nz * Lz, dls = nx * hx + ny * hy + nz * hz;
            dld = dld < 0 ? 0 : (dld > 1 ? 1 : dld);
            dls = dls < 0 ? 0 : (dls > 1 ? 1 : dls);
            dls = pow(dls, (double)m->shininess);
            for (int k = 0; k < 3; k++) {
                dLd[k] += dld * col[k] * dI;
                dLs[k] += dls * l->color[k] * dI;
            }
        }
    }
    float r = LaR + ((float)m->diffuse[0] * LdR / 255.0f) + ((float)m->specular[0] * LsR / 255.0f);
    float g = LaG + ((float)m->diffuse[1] * LdG / 255.0f) + ((float)m->specular[1] * LsG / 255.0f);
    float b = LaB + ((float)m->diffuse[2] * LdB / 255.0f) + ((float)m->specular[2] * LsB / 255.0f);
    if (lin) lin[0] = r, lin[1] = g, lin[2] = b;
    if (lin64)
        for (int k = 0; k < 3; k++)
            lin64[k] = dLa[k] + (double)m->diffuse[k] * dLd[k] / 255.0 + (double)m->specular[k] * dLs[k] / 255.0;
    out[0] = gf2u8(gclamp(ground(r), 0, 0xff));
    out[1] = gf2u8(gclamp(ground(g), 0, 0xff));
    out[2] = gf2u8(gclamp(ground(b), 0, 0xff));
    out[3] = gf2u8(gclamp((float)col[3], 0, 0xff));
    return 1;
}

/* material/ao.go:26-28: the eight march directions. */
static void ao_dirs(float out[16]) {
    int k = 0;
    for (float a = 0.0f; a < (float)(2 * M_PI - 1e-4); a += (float)(M_PI / 4)) {
        out[2 * k] = (float)cos((double)a);
        out[2 * k + 1] = (float)sin((double)a);
        k++;
    }
}
/* material/ao.go:40-73 */
static float max_elevation_angle(ctx_t* c, int px, int py, float dx, float dy) {
    v4 p = V4((float)px, (float)py, 0, 1);
    v4 dir = V4(dx, dy, 0, 0);
    float maxangle = 0.0f;
    for (float t = 0.0f; t < 100; t += 1) {
        v4 cur = add4(p, scale4(dir, t, t, 1, 1));
        int64_t cx = gf2i(cur.x), cy = gf2i(cur.y);
        if (!in_buf(c, cx, cy)) return maxangle;
        float distance = len4(sub4(p, cur));
        if (distance < 1) continue;
        const frag_t* sh = fat(c, (int)cx, (int)cy);
        const frag_t* tr = fat(c, (int)gf2i(p.x), (int)gf2i(p.y));
        float sd = sh->ok ? sh->depth : -1, td = tr->ok ? tr->depth : -1;
        float elevation = sd - td;
        maxangle = gmax2(maxangle, (float)atan((double)(elevation / distance)));
    }
    return maxangle;
}
/* material/ao.go:20-38 */
static void ao_shade(ctx_t* c, int fx, int fy, const uint8_t col[4], const prh_material* mat, uint8_t out[4],
                     float* total_out) {
    if (!mat || !mat->ambient_occlusion) {
        memcpy(out, col, 4);
        return;
    }
    float total = 0.0f;
    for (int k = 0; k < 8; k++)
        total += (float)(M_PI / 2) - max_elevation_angle(c, fx, fy, c->ao_dir[2 * k], c->ao_dir[2 * k + 1]);
    total /= (float)((M_PI / 2) * 8);
    if (total_out) *total_out = total;
    total = gpow(total, 10000);
    out[0] = gf2u8(total * (float)col[0]);
    out[1] = gf2u8(total * (float)col[1]);
    out[2] = gf2u8(total * (float)col[2]);
    out[3] = col[3];
}

/* render/raster.go:324-359, called from DrawFragment (render/raster_screen.go:75-87)
 * with frag = the pixel's stored fragment. */
static void shade(ctx_t* c, int x, int y) {
    frag_t* fr = fat(c, x, y); /* DrawFragment's info (copy written back below) */
    frag_t frag = *fr;
    const frag_t* info = fat(c, frag.x, frag.y);
    uint8_t col[4];
    size_t lid = (size_t)y * c->W + x;
    if (!info->ok) {
        memcpy(col, c->o->background, 4);
    } else {
        memcpy(col, info->col, 4);
        const prh_material* mat = mat_at(c->s, frag.matid);
        if (mat) {
            float lin[3];
            double lin64[3];
            /* the texture colour: Query at the fragment's UV and mip level, or
             * the uniform colour */
            uint8_t tcol[4];
            memcpy(tcol, mat->texture, 4);
            if (mat->texture_id >= 0 && mat->texture_id < c->n_tex)
                tex_fragment(&c->tex[mat->texture_id], info->du, info->dv, info->u, info->v, tcol);
            int lit = fragment_shader(mat, tcol, info->nor, info->facenor, info->wpos, V3a(c->cam->position),
                                      c->s->n_lights, c->s->lights, c->s->n_ambient, c->s->ambient_intensity, col,
                                      lin, lin64);
            if (lit && c->linear) {
                c->linear[3 * lid] = lin[0];
                c->linear[3 * lid + 1] = lin[1];
                c->linear[3 * lid + 2] = lin[2];
                for (int k = 0; k < 3; k++) {
                    double e = fabs((double)lin[k] - lin64[k]);
                    if (e > c->shade_err64) c->shade_err64 = e;
                }
            }
            if (c->o->shadow_map && mat->receive_shadow) {
                float vis = 0.0f;
                for (int i = 0; i < c->s->n_lights; i++)
                    if (shading_visibility(c, i, info)) vis++;
                if (c->shadow_count && frag.ok) c->shadow_count[lid] = (uint8_t)vis;
                float w = gpow(0.5f, vis);
                for (int k = 0; k < 3; k++) col[k] = gf2u8((float)col[k] * w);
            }
        }
        uint8_t outc[4];
        ao_shade(c, frag.x, frag.y, col, mat_at(c->s, frag.matid), outc, c->ao_total ? &c->ao_total[lid] : NULL);
        memcpy(col, outc, 4);
    }
    memcpy(fr->col, col, 4);
}

/* ----------------------------------------------------------------- render */
int oracle_render(const prh_scene_desc* s, const prh_camera* cam, const prh_options* o, oracle_out* out) {
    if (!s || !cam || !o || !out) return PRH_E_INVALID;
    int msaa = o->msaa > 0 ? o->msaa : 1;
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.s = s;
    c.cam = cam;
    c.o = o;
    c.W = o->width * msaa;
    c.H = o->height * msaa;
    c.perspect = cam->kind == PRH_CAMERA_PERSPECTIVE;
    c.SW = o->shadow_size > 0 ? o->shadow_size : c.W;
    c.SH = o->shadow_size > 0 ? o->shadow_size : c.H;
    size_t np = (size_t)c.W * c.H;
    c.f = calloc(np, sizeof(frag_t));
    if (!c.f) return PRH_E_NOMEM;
    c.workers = g_workers;
    if (c.workers > 1) {
        c.locks = calloc(np, 1);
        if (!c.locks) {
            free(c.f);
            return PRH_E_NOMEM;
        }
    }
    if (s->n_textures > 0) {
        c.tex = calloc((size_t)s->n_textures, sizeof(tex_t));
        int bad = !c.tex;
        for (int i = 0; i < s->n_textures && !bad; i++, c.n_tex++) bad = tex_build(&s->textures[i], &c.tex[i]) != 0;
        if (bad) {
            for (int i = 0; i < c.n_tex; i++) tex_free(&c.tex[i]);
            free(c.tex);
            free(c.f);
            return PRH_E_NOMEM;
        }
    }
    c.linear = out->linear;
    if (c.linear)
        for (size_t i = 0; i < 3 * np; i++) c.linear[i] = NAN;
    c.shadow_count = out->shadow_count;
    if (c.shadow_count) memset(c.shadow_count, 0, np);
    c.ao_total = out->ao_total;
    if (c.ao_total)
        for (size_t i = 0; i < np; i++) c.ao_total[i] = NAN;
    ao_dirs(c.ao_dir);
    c.view = view_matrix(V3a(cam->position), V3a(cam->target), V3a(cam->up));
    c.proj = cam_proj(cam);
    c.vp = viewport_matrix((float)c.W, (float)c.H);
    c.viewinv = inv(&c.view);
    c.projinv = inv(&c.proj);
    c.vpinv = inv(&c.vp);
    m4 t0 = mulm(&c.viewinv, &c.projinv);
    c.s2w = mulm(&t0, &c.vpinv);

    if (o->shadow_map) {
        init_shadow_maps(&c);
        for (int i = 0; i < s->n_lights; i++) pass_shadows(&c, i);
    }
    forward_pass(&c);
    for (int y = 0; y < c.H; y++)
        for (int x = 0; x < c.W; x++) {
            const frag_t* f = fat(&c, x, y);
            size_t i = (size_t)y * c.W + x;
            if (out->vis_order) out->vis_order[i] = f->ok ? f->order : 0xFFFFFFFFu;
            if (out->depth) out->depth[i] = f->ok ? f->depth : -1.0f;
        }
    /* deferred: DrawFragments order j = 0..n-1, x = j%w, y = j/w; pixel (0,0)
     * is first, so every later pixel may read its shaded colour. */
    shade(&c, 0, 0);
    const uint8_t* mask = out->shade_mask;
#ifdef _OPENMP
    /* (the deferred pass: on OpenMP's threads in the parity order, on one
     * in Mode A, on `workers` in Mode B) */
    int nthr = c.workers == 0 ? omp_get_max_threads() : c.workers;
#pragma omp parallel num_threads(nthr)
    {
        ctx_t cl = c;
        cl.shade_err64 = 0;
#pragma omp for schedule(dynamic, 256)
        for (long j = 1; j < (long)np; j++) {
            int x = (int)(j % c.W), y = (int)(j / c.W);
            if (mask && !mask[j]) continue;
            shade(&cl, x, y);
        }
#pragma omp critical
        if (cl.shade_err64 > c.shade_err64) c.shade_err64 = cl.shade_err64;
    }
#else
    for (size_t j = 1; j < np; j++) {
        int x = (int)(j % c.W), y = (int)(j / c.W);
        if (mask && !mask[j]) continue;
        shade(&c, x, y);
    }
#endif
    /* passAntialiasing (render/raster.go:361-378): gamma on the
     * MSAA-sized buffer, then Resize to the requested size */
    uint8_t* big = msaa > 1 && out->image ? malloc(np * 4) : out->image;
    int rc = PRH_OK;
    if (msaa > 1 && out->image && !big) rc = PRH_E_NOMEM;
    for (int y = 0; y < c.H && rc == PRH_OK; y++)
        for (int x = 0; x < c.W; x++) {
            size_t i = (size_t)y * c.W + x;
            const frag_t* f = fat(&c, x, y);
            uint8_t col[4];
            memcpy(col, f->col, 4);
            if (mask && !mask[i] && i != 0) memset(col, 0, 4);
            if (out->pregamma) memcpy(out->pregamma + 4 * i, col, 4);
            if (o->gamma_correct) /* shader/gamma.go:13-18 */
                for (int k = 0; k < 3; k++) col[k] = gamma_u8(col[k]);
            if (big) memcpy(big + 4 * ((size_t)(c.H - 1 - y) * c.W + x), col, 4);
        }
    if (rc == PRH_OK && msaa > 1 && out->image) {
        if (resize_image(big, c.W, c.H, out->image, o->width, o->height)) rc = PRH_E_NOMEM;
        free(big);
    }
    if (out->shadow_maps && c.smaps)
        memcpy(out->shadow_maps, c.smaps, sizeof(float) * (size_t)s->n_lights * c.SW * c.SH);
    if (out->shade_err64) *out->shade_err64 = c.shade_err64;
    if (out->light_matrices && c.lview)
        for (int i = 0; i < s->n_lights; i++) {
            memcpy(out->light_matrices + 32 * (size_t)i, c.lview[i].a, 64);
            memcpy(out->light_matrices + 32 * (size_t)i + 16, c.lproj[i].a, 64);
        }
    free(c.f);
    free(c.locks);
    for (int i = 0; i < c.n_tex; i++) tex_free(&c.tex[i]);
    free(c.tex);
    free(c.lview);
    free(c.lproj);
    free(c.lactive);
    free(c.smaps);
    return rc;
}

/* --------------------------------------------------- differentiable shading
 * The FP64 reference of libprh's adjoint (an extension: polyred has no
 * derivative path).  At fixed visibility, each covered pixel with a material
 * is shaded with the Blinn-Phong sums of shader/blinn_cpu.go:24-106 (before
 * Round), except that its albedo is the perspective-correct interpolation of
 * per-vertex albedo parameters (where the reference reads the uniform
 * texture colour) and its normal the unit interpolation of per-vertex normal
 * parameters; Kd, Ks and the shininess are real-valued.  A clipped piece
 * takes its parent triangle's vertices.  The geometry (barycentrics at the
 * pixel centre, perspective weights, world position) is float32 as in the
 * forward pass; the shading, the loss sum((lin - target)^2) and the
 * gradients are float64. */
/* parameter layout: albedo [Vt][3] per transformed vertex; normal [Vt][3] or
 * NULL (n.Apply(Normal) of the mesh normals); kd, ks [M][3]; shininess [M];
 * light_pos [L][3]; light_int [L]. */
static void dunit3(const double v[3], double out[3], double* len) {
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    *len = l;
    for (int k = 0; k < 3; k++) out[k] = v[k] / l;
}
/* d(v/|v|) applied to an upstream gradient: (g - u (u.g)) / |v| */
static void dunit3_back(const double u[3], double len, const double g[3], double out[3]) {
    double ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];
    for (int k = 0; k < 3; k++) out[k] = (g[k] - u[k] * ug) / len;
}

int oracle_diff(const prh_scene_desc* s, const prh_camera* cam, const prh_options* o, const uint32_t* vis_order,
                const oracle_diff_params* p, const double* target, double* linear, oracle_diff_grads* g,
                double* loss) {
    if (!s || !cam || !o || !vis_order || !p || !loss) return PRH_E_INVALID;
    const int msaa = o->msaa > 0 ? o->msaa : 1, W = o->width * msaa, H = o->height * msaa;
    const int perspect = cam->kind == PRH_CAMERA_PERSPECTIVE;
    /* per transformed vertex: screen xy + recipw, transformed normal, back-transformed position */
    int64_t nv = 0, nt = 0;
    for (int ob = 0; ob < s->n_objects; ob++) {
        int m = s->obj_mesh[ob];
        nv += s->mesh_vtx_off[m + 1] - s->mesh_vtx_off[m];
        nt += s->mesh_tri_off[m + 1] - s->mesh_tri_off[m];
    }
    v4* scr = malloc(sizeof(v4) * (size_t)nv);
    v4* tnor = malloc(sizeof(v4) * (size_t)nv);
    v4* wp = malloc(sizeof(v4) * (size_t)nv);
    int64_t* tri_start = malloc(sizeof(int64_t) * (size_t)(s->n_objects + 1));
    int64_t* vtx_start = malloc(sizeof(int64_t) * (size_t)(s->n_objects + 1));
    if (!scr || !tnor || !wp || !tri_start || !vtx_start) {
        free(scr), free(tnor), free(wp), free(tri_start), free(vtx_start);
        return PRH_E_NOMEM;
    }
    m4 view = view_matrix(V3a(cam->position), V3a(cam->target), V3a(cam->up)), proj = cam_proj(cam);
    m4 vp = viewport_matrix((float)W, (float)H);
    m4 vinv = inv(&view), pinv = inv(&proj), vpinv = inv(&vp), pv = mulm(&proj, &view);
    tri_start[0] = vtx_start[0] = 0;
    for (int ob = 0; ob < s->n_objects; ob++) {
        int m = s->obj_mesh[ob];
        m4 model = obj_model(s, ob), mi = inv(&model), normal = transpose(&mi), trans = mulm(&pv, &model);
        int64_t base = vtx_start[ob];
        for (int v = s->mesh_vtx_off[m]; v < s->mesh_vtx_off[m + 1]; v++) {
            vert_t in;
            load_vert(s, v, &in);
            v4 q = mulv(&trans, in.pos);
            float rw = perspect ? -1 / q.w : 1.0f;
            v4 sp = pos4(apply4(q, &vp));
            int64_t id = base + (v - s->mesh_vtx_off[m]);
            scr[id] = V4(sp.x, sp.y, sp.z, rw);
            tnor[id] = apply4(in.nor, &normal);
            wp[id] = apply4(apply4(apply4(sp, &vpinv), &pinv), &vinv);
        }
        vtx_start[ob + 1] = base + (s->mesh_vtx_off[m + 1] - s->mesh_vtx_off[m]);
        tri_start[ob + 1] = tri_start[ob] + (s->mesh_tri_off[m + 1] - s->mesh_tri_off[m]);
    }
    double amb = 0;
    for (int e = 0; e < s->n_ambient; e++) amb += s->ambient_intensity[e];
    const double camp[3] = {cam->position[0], cam->position[1], cam->position[2]};
    double L2 = 0;
    for (int64_t i = 0; i < (int64_t)W * H; i++) {
        const int x = (int)(i % W), y = (int)(i / W);
        if (linear) linear[3 * i] = linear[3 * i + 1] = linear[3 * i + 2] = NAN;
        const uint32_t ord = vis_order[i];
        if (ord == 0xFFFFFFFFu) continue;
        const int64_t gt = ord >> 3;
        int ob = 0;
        while (tri_start[ob + 1] <= gt) ob++;
        const int mesh = s->obj_mesh[ob];
        const int t = s->mesh_tri_off[mesh] + (int)(gt - tri_start[ob]);
        int64_t mat = s->tri_mat[t];
        if (mat >= 0) mat += s->obj_mat_base[ob];
        if (mat < 0 || mat >= s->n_materials) continue;
        const prh_material* mt = &s->materials[mat];
        int64_t id[3];
        for (int k = 0; k < 3; k++) id[k] = vtx_start[ob] + s->tri_v[3 * t + k];
        float bc[3];
        barycoord((float)x + 0.5f, (float)y + 0.5f, scr[id[0]].x, scr[id[0]].y, scr[id[1]].x, scr[id[1]].y,
                  scr[id[2]].x, scr[id[2]].y, bc);
        float wc[3] = {scr[id[0]].w * bc[0], scr[id[1]].w * bc[1], scr[id[2]].w * bc[2]};
        float norm = perspect ? 1 / (wc[0] + wc[1] + wc[2]) : 1.0f;
        double w[3], a[3] = {0, 0, 0}, nraw[3] = {0, 0, 0}, xw[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++) {
            w[k] = (double)(wc[k] * norm);
            const double* A = p->albedo + 3 * id[k];
            double N[3];
            if (p->normal) N[0] = p->normal[3 * id[k]], N[1] = p->normal[3 * id[k] + 1], N[2] = p->normal[3 * id[k] + 2];
            else N[0] = tnor[id[k]].x, N[1] = tnor[id[k]].y, N[2] = tnor[id[k]].z;
            for (int c2 = 0; c2 < 3; c2++) {
                a[c2] += w[k] * A[c2];
                nraw[c2] += (double)bc[k] * N[c2];
            }
        }
        v4 xf = interp_wpos(bc, wp[id[0]], wp[id[1]], wp[id[2]]);
        xw[0] = xf.x, xw[1] = xf.y, xw[2] = xf.z;
        double n[3], nl;
        if (mt->flat_shading) {
            v4 m1 = wp[id[0]], m2 = wp[id[1]], m3 = wp[id[2]];
            v4 fN = unit4(cross4(sub4(m2, m1), sub4(m3, m1)));
            n[0] = fN.x, n[1] = fN.y, n[2] = fN.z, nl = 1;
        } else {
            dunit3(nraw, n, &nl);
        }
        double vv[3] = {camp[0] - xw[0], camp[1] - xw[1], camp[2] - xw[2]}, V[3], vl;
        dunit3(vv, V, &vl);
        const double* kd = p->kd + 3 * mat;
        const double* ks = p->ks + 3 * mat;
        const double sh = p->shininess[mat];
        enum { MAXL = 64 };
        double Ld[MAXL], Ls[MAXL], I[MAXL], Lv[MAXL][3], Hv[MAXL][3], hl[MAXL], dl[MAXL], dv[MAXL][3], nL[MAXL],
            nH[MAXL];
        const int NL = s->n_lights < MAXL ? s->n_lights : MAXL;
        double sLd = 0, sLs[3] = {0, 0, 0};
        for (int li = 0; li < NL; li++) {
            const prh_light* l = &s->lights[li];
            if (l->kind == PRH_LIGHT_POINT) {
                for (int k = 0; k < 3; k++) dv[li][k] = p->light_pos[3 * li + k] - xw[k];
                dunit3(dv[li], Lv[li], &dl[li]);
                I[li] = p->light_int[li] / dl[li];
            } else {
                for (int k = 0; k < 3; k++) Lv[li][k] = -(double)l->direction[k];
                dl[li] = 1;
                I[li] = p->light_int[li];
            }
            double hv[3] = {Lv[li][0] + V[0], Lv[li][1] + V[1], Lv[li][2] + V[2]};
            dunit3(hv, Hv[li], &hl[li]);
            nL[li] = n[0] * Lv[li][0] + n[1] * Lv[li][1] + n[2] * Lv[li][2];
            nH[li] = n[0] * Hv[li][0] + n[1] * Hv[li][1] + n[2] * Hv[li][2];
            Ld[li] = nL[li] < 0 ? 0 : (nL[li] > 1 ? 1 : nL[li]);
            const double cH = nH[li] < 0 ? 0 : (nH[li] > 1 ? 1 : nH[li]);
            Ls[li] = pow(cH, sh);
            sLd += Ld[li] * I[li];
            for (int c2 = 0; c2 < 3; c2++) sLs[c2] += Ls[li] * (double)l->color[c2] * I[li];
        }
        double lin[3], gl[3];
        /* fixed shadow visibility: the reference halves the shaded colour
         * per reporting light (render/raster.go:339-347) */
        const double sw = p->shadow_count ? ldexp(1.0, -(int)p->shadow_count[i]) : 1.0;
        for (int c2 = 0; c2 < 3; c2++) {
            lin[c2] = sw * (amb * a[c2] + kd[c2] * sLd * a[c2] / 255.0 + ks[c2] * sLs[c2] / 255.0);
            if (linear) linear[3 * i + c2] = lin[c2];
            const double d = target ? lin[c2] - target[3 * i + c2] : 0;
            L2 += d * d;
            gl[c2] = 2 * d * sw;
        }
        if (!g || !target) continue;
        /* backward */
        double ga[3], gn[3] = {0, 0, 0};
        for (int c2 = 0; c2 < 3; c2++) {
            if (g->kd) g->kd[3 * mat + c2] += gl[c2] * sLd * a[c2] / 255.0;
            if (g->ks) g->ks[3 * mat + c2] += gl[c2] * sLs[c2] / 255.0;
            ga[c2] = gl[c2] * (amb + kd[c2] * sLd / 255.0);
        }
        for (int li = 0; li < NL; li++) {
            const prh_light* l = &s->lights[li];
            double gLd = 0, gLs = 0, gI = 0;
            for (int c2 = 0; c2 < 3; c2++) {
                gLd += gl[c2] * kd[c2] * a[c2] * I[li] / 255.0;
                gLs += gl[c2] * ks[c2] * (double)l->color[c2] * I[li] / 255.0;
                gI += gl[c2] * (kd[c2] * Ld[li] * a[c2] + ks[c2] * Ls[li] * (double)l->color[c2]) / 255.0;
            }
            const double dLd = (nL[li] > 0 && nL[li] < 1) ? 1 : 0;
            double dLs = 0;
            if (nH[li] > 0 && nH[li] < 1) {
                dLs = sh * pow(nH[li], sh - 1);
                if (g->shininess) g->shininess[mat] += gLs * Ls[li] * log(nH[li]);
            }
            double gL[3], gH[3], ghv[3];
            for (int k = 0; k < 3; k++) {
                gn[k] += gLd * dLd * Lv[li][k] + gLs * dLs * Hv[li][k];
                gL[k] = gLd * dLd * n[k];
                gH[k] = gLs * dLs * n[k];
            }
            dunit3_back(Hv[li], hl[li], gH, ghv);
            for (int k = 0; k < 3; k++) gL[k] += ghv[k];
            if (l->kind == PRH_LIGHT_POINT) {
                double gd[3];
                dunit3_back(Lv[li], dl[li], gL, gd);
                const double i0 = p->light_int[li], d3 = dl[li] * dl[li] * dl[li];
                for (int k = 0; k < 3; k++) gd[k] += gI * (-i0 * dv[li][k] / d3);
                if (g->light_pos)
                    for (int k = 0; k < 3; k++) g->light_pos[3 * li + k] += gd[k];
                if (g->light_int) g->light_int[li] += gI / dl[li];
            } else if (g->light_int) {
                g->light_int[li] += gI;
            }
        }
        double gnr[3] = {0, 0, 0};
        if (!mt->flat_shading) dunit3_back(n, nl, gn, gnr);
        for (int k = 0; k < 3; k++)
            for (int c2 = 0; c2 < 3; c2++) {
                if (g->albedo) g->albedo[3 * id[k] + c2] += w[k] * ga[c2];
                if (g->normal) g->normal[3 * id[k] + c2] += (double)bc[k] * gnr[c2];
            }
    }
    *loss = L2;
    free(scr), free(tnor), free(wp), free(tri_start), free(vtx_start);
    return PRH_OK;
}

/* ----------------------------------------------------- known-answer hooks */
void oracle_barycoord(const float p[2], const float a[2], const float b[2], const float c[2], float out[3]) {
    barycoord(p[0], p[1], a[0], a[1], b[0], b[1], c[0], c[1], out);
}
void oracle_viewport_matrix(float w, float h, float out[16]) { memcpy(out, viewport_matrix(w, h).a, 64); }
void oracle_view_matrix(const float pos[3], const float target[3], const float up[3], float out[16]) {
    memcpy(out, view_matrix(V3a(pos), V3a(target), V3a(up)).a, 64);
}
void oracle_proj_matrix(const prh_camera* cam, float out[16]) { memcpy(out, cam_proj(cam).a, 64); }
void oracle_mat4_inv(const float m[16], float out[16]) {
    m4 a;
    memcpy(a.a, m, 64);
    memcpy(out, inv(&a).a, 64);
}
void oracle_mat4_mulm(const float a[16], const float b[16], float out[16]) {
    m4 x, y;
    memcpy(x.a, a, 64);
    memcpy(y.a, b, 64);
    memcpy(out, mulm(&x, &y).a, 64);
}
int oracle_aabb_intersect(const float min1[3], const float max1[3], const float min2[3], const float max2[3]) {
    aabb_t a = {V3a(min1), V3a(max1)}, b = {V3a(min2), V3a(max2)};
    return aabb_intersect(&a, &b);
}
int oracle_aabb_contains(const float mn[3], const float mx[3], int n, const float* pts) {
    aabb_t a = {V3a(mn), V3a(mx)};
    v3 vs[64];
    if (n > 64) return -1;
    for (int i = 0; i < n; i++) vs[i] = V3a(pts + 3 * i);
    return aabb_contains(&a, n, vs);
}
void oracle_aabb_new(int n, const float* pts, float mn[3], float mx[3]) {
    v3 vs[64];
    if (n > 64) n = 64;
    for (int i = 0; i < n; i++) vs[i] = V3a(pts + 3 * i);
    aabb_t b = aabb_new(n, vs);
    mn[0] = b.mn.x, mn[1] = b.mn.y, mn[2] = b.mn.z;
    mx[0] = b.mx.x, mx[1] = b.mx.y, mx[2] = b.mx.z;
}
int oracle_tri_is_valid(const float p1[4], const float p2[4], const float p3[4]) {
    return tri_is_valid(V4(p1[0], p1[1], p1[2], p1[3]), V4(p2[0], p2[1], p2[2], p2[3]),
                        V4(p3[0], p3[1], p3[2], p3[3]));
}
void oracle_interp_worldpos(const float bc[3], const float m1[4], const float m2[4], const float m3[4],
                            float out[4]) {
    v4 r = interp_wpos(bc, V4(m1[0], m1[1], m1[2], m1[3]), V4(m2[0], m2[1], m2[2], m2[3]),
                       V4(m3[0], m3[1], m3[2], m3[3]));
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
int oracle_mat_at(int32_t n, int64_t id) { return id >= 0 && id < n ? (int)id : -1; }
float oracle_srgb_from_linear(float v) { return from_linear_srgb(v); }
float oracle_linear_from_srgb(float v) { return from_srgb_linear(v); }
uint8_t oracle_gamma_u8(uint8_t c) { return gamma_u8(c); }
uint8_t oracle_shadow_quant(uint8_t c, int occluded) { return gf2u8((float)c * gpow(0.5f, (float)occluded)); }
void oracle_fragment_shader(const prh_material* m, const float nor[4], const float facenor[4], const float wpos[4],
                            const float cam[3], int n_lights, const prh_light* lights, int n_amb, const float* amb,
                            uint8_t out[4], float lin[3]) {
    float l[3] = {NAN, NAN, NAN};
    fragment_shader(m, m->texture, V4(nor[0], nor[1], nor[2], nor[3]), V4(facenor[0], facenor[1], facenor[2], facenor[3]),
                    V4(wpos[0], wpos[1], wpos[2], wpos[3]), V3a(cam), n_lights, lights, n_amb, amb, out, l, NULL);
    if (lin) memcpy(lin, l, sizeof l);
}
void oracle_ao_dirs(float out[16]) { ao_dirs(out); }

int oracle_ao_march(const float* depth, const uint8_t* ok, int W, int H, int px, int py, float angles[8],
                    float* total, float* factor) {
    if (!depth || !ok || W <= 0 || H <= 0 || px < 0 || px >= W || py < 0 || py >= H) return PRH_E_INVALID;
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.W = W;
    c.H = H;
    c.f = calloc((size_t)W * H, sizeof(frag_t));
    if (!c.f) return PRH_E_NOMEM;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            frag_t* f = fat(&c, x, y);
            f->ok = ok[(size_t)y * W + x] != 0;
            f->depth = depth[(size_t)y * W + x];
            f->x = x;
            f->y = y;
        }
    ao_dirs(c.ao_dir);
    for (int k = 0; k < 8 && angles; k++) angles[k] = max_elevation_angle(&c, px, py, c.ao_dir[2 * k], c.ao_dir[2 * k + 1]);
    prh_material m;
    memset(&m, 0, sizeof m);
    m.ambient_occlusion = 1;
    const uint8_t white[4] = {255, 255, 255, 255};
    uint8_t out[4];
    float t = NAN;
    ao_shade(&c, px, py, white, &m, out, &t);
    if (total) *total = t;
    if (factor) *factor = gpow(t, 10000);
    free(c.f);
    return PRH_OK;
}

int oracle_shadow_lookup(const float s2w[16], const float lview[16], const float lproj[16], int SW, int SH,
                         const float* map, float x, float y, float depth) {
    m4 a, b, d;
    memcpy(a.a, s2w, 64);
    memcpy(b.a, lview, 64);
    memcpy(d.a, lproj, 64);
    return shadow_lookup(&a, &b, &d, SW, SH, map, x, y, depth);
}
