// runtime.hip — libprh's host runtime and C ABI (include/prh.h).
//
// Host-side work mirrors what render.NewRenderer / (*Renderer).Render do on
// the CPU before touching pixels (render/raster.go:84-143,155-199,226-315,
// render/shadow.go:33-90): camera, projection and viewport matrices and their
// inverses, per-object Model^-1^T, the orthographic light cameras fitted to
// the camera frustum, the sRGB table.  Everything per pixel or per triangle
// runs in kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/prh.h"
#include "kernels.hpp"

using namespace gm;
using namespace prh;

// prh_diff_backward_noalloc's working memory: one array per GPU, which the
// code object's loader places when the library is first used on that GPU.
// Host code never passes `scratch` itself to a kernel or a copy (on the host
// it names a stand-in, not the device array): prh_open looks up its device
// address once per GPU and every later use goes through that pointer.
static __device__ __align__(256) uint8_t scratch[PRH_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? PRH_E_NOMEM : PRH_E_HIP,                       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    } while (0)

std::mutex g_scratch_mu;
std::vector<uint8_t*> g_scratch_dev;  // per device ordinal: scratch's device address, or null

// With `device` current: the device address of `scratch` on it, looked up on
// the first call for that device.
int scratch_setup(int device, uint8_t** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)device >= g_scratch_dev.size()) g_scratch_dev.resize((size_t)device + 1, nullptr);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return PRH_E_HIP;
        }
        g_scratch_dev[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch_dev[device];
    return PRH_OK;
}

// ---------------------------------------------------------- host matrices
M4 mat(const float* a) {
    M4 m;
    memcpy(m.a, a, sizeof m.a);
    return m;
}
M4 mulm(const M4& m, const M4& n) {  // math/mat4.go:201-219
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.a[i * 4 + j] = m.a[i * 4] * n.a[j] + m.a[i * 4 + 1] * n.a[4 + j] + m.a[i * 4 + 2] * n.a[8 + j] +
                             m.a[i * 4 + 3] * n.a[12 + j];
    return r;
}
M4 transpose(const M4& m) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.a[i * 4 + j] = m.a[j * 4 + i];
    return r;
}
// Mat4.Det / Mat4.Inv (math/mat4.go:233-283).  The float32 result depends on
// the order of the products and sums, so the terms are listed in the
// reference's order: a sign and the element indices of each product.
const char* const kDetTerms =
    "+0,5,10,15 -0,5,11,14 +0,6,11,13 -0,6,9,15 +0,7,9,14 -0,7,10,13 -1,6,11,12 +1,6,8,15 "
    "-1,7,8,14 +1,7,10,12 -1,4,10,15 +1,4,11,14 +2,7,8,13 -2,7,9,12 +2,4,9,15 -2,4,11,13 "
    "+2,5,11,12 -2,5,8,15 -3,4,9,14 +3,4,10,13 -3,5,10,12 +3,5,8,14 -3,6,8,13 +3,6,9,12";
const char* const kInvTerms[16] = {
    "+6,11,13 -7,10,13 +7,9,14 -5,11,14 -6,9,15 +5,10,15", "+3,10,13 -2,11,13 -3,9,14 +1,11,14 +2,9,15 -1,10,15",
    "+2,7,13 -3,6,13 +3,5,14 -1,7,14 -2,5,15 +1,6,15",     "+3,6,9 -2,7,9 -3,5,10 +1,7,10 +2,5,11 -1,6,11",
    "+7,10,12 -6,11,12 -7,8,14 +4,11,14 +6,8,15 -4,10,15", "+2,11,12 -3,10,12 +3,8,14 -0,11,14 -2,8,15 +0,10,15",
    "+3,6,12 -2,7,12 -3,4,14 +0,7,14 +2,4,15 -0,6,15",     "+2,7,8 -3,6,8 +3,4,10 -0,7,10 -2,4,11 +0,6,11",
    "+5,11,12 -7,9,12 +7,8,13 -4,11,13 -5,8,15 +4,9,15",   "+3,9,12 -1,11,12 -3,8,13 +0,11,13 +1,8,15 -0,9,15",
    "+1,7,12 -3,5,12 +3,4,13 -0,7,13 -1,4,15 +0,5,15",     "+3,5,8 -1,7,8 -3,4,9 +0,7,9 +1,4,11 -0,5,11",
    "+6,9,12 -5,10,12 -6,8,13 +4,10,13 +5,8,14 -4,9,14",   "+1,10,12 -2,9,12 +2,8,13 -0,10,13 -1,8,14 +0,9,14",
    "+2,5,12 -1,6,12 -2,4,13 +0,6,13 +1,4,14 -0,5,14",     "+1,6,8 -2,5,8 +2,4,9 -0,6,9 -1,4,10 +0,5,10"};

// Evaluates "+i,j,k ..." left to right: acc = p0; acc = acc +/- p_t.
float eval_terms(const float* a, const char* terms) {
    float acc = 0;
    bool first = true;
    const char* p = terms;
    while (*p) {
        while (*p == ' ') p++;
        if (!*p) break;
        char sign = *p++;
        float prod = 0;
        bool firstf = true;
        for (;;) {
            int idx = (int)strtol(p, (char**)&p, 10);
            prod = firstf ? a[idx] : prod * a[idx];
            firstf = false;
            if (*p == ',') {
                p++;
                continue;
            }
            break;
        }
        acc = first ? prod : (sign == '+' ? acc + prod : acc - prod);
        first = false;
    }
    return acc;
}
float det(const M4& m) { return eval_terms(m.a, kDetTerms); }
M4 inv(const M4& m) {
    float dinv = 1 / det(m);
    M4 r;
    for (int e = 0; e < 16; e++) r.a[e] = dinv * eval_terms(m.a, kInvTerms[e]);
    return r;
}
M4 viewport_matrix(float w, float h) {  // math/math.go:270-277
    M4 r = {{w / 2, 0, 0, w / 2, 0, h / 2, 0, h / 2, 0, 0, 1, 0, 0, 0, 0, 1}};
    return r;
}
V3 v3a(const float* p) { return v3(p[0], p[1], p[2]); }
M4 view_matrix(V3 pos, V3 target, V3 up) {  // camera/camera.go:42-53
    V3 l = unit(sub(target, pos));
    V3 lxu = unit(cross(l, up));
    V3 u = unit(cross(lxu, l));
    M4 r = {{lxu.x, lxu.y, lxu.z, -dot(lxu, pos), u.x, u.y, u.z, -dot(u, pos), -l.x, -l.y, -l.z, dot(l, pos), 0, 0, 0,
             1}};
    return r;
}
M4 persp_proj(float fov_deg, float aspect, float n, float f) {  // camera/perspective.go:100-111
    float fov = (fov_deg * (float)M_PI) / 180;
    float t = (float)std::tan((double)(fov / 2));
    M4 r = {{-1 / (aspect * t), 0, 0, 0, 0, -1 / t, 0, 0, 0, 0, (n + f) / (n - f), (2 * n * f) / (n - f), 0, 0, 1, 0}};
    return r;
}
M4 ortho_proj(float l, float r_, float b, float t, float n, float f) {  // camera/orthographic.go:106-117
    M4 r = {{2 / (r_ - l), 0, 0, (l + r_) / (l - r_), 0, 2 / (t - b), 0, (b + t) / (b - t), 0, 0, 2 / (n - f),
             (f + n) / (f - n), 0, 0, 0, 1}};
    return r;
}
M4 cam_proj(const prh_camera& c) {
    if (c.kind == PRH_CAMERA_PERSPECTIVE) return persp_proj(c.fov, c.aspect, c.near_, c.far_);
    return ortho_proj(c.left, c.right, c.bottom, c.top, c.near_, c.far_);
}
// Triangle.IsValid (geometry/primitive/triangle.go:63-80), model space.
bool tri_is_valid(V4 p1, V4 p2, V4 p3) {
    const float eps = 1e-7f;
    V4 a = sub(p2, p1), b = sub(p3, p1);
    if (approx_eq(a.x, 0, eps) && approx_eq(a.y, 0, eps) && approx_eq(a.z, 0, eps)) return false;
    if (approx_eq(b.x, 0, eps) && approx_eq(b.y, 0, eps) && approx_eq(b.z, 0, eps)) return false;
    float d = dot(a, b) / (len(a) * len(b));
    return !approx_eq(d, 1, eps) && !approx_eq(d, -1, eps);
}
// color/srgb.go:57-93 and shader/gamma.go:13-18 for each 8-bit value.
void gamma_table(uint8_t out[256]) {
    double lut[1025];
    for (int i = 0; i < 1024; i++) {
        double v = (double)i / 1024;
        lut[i] = v <= 0.0031308 ? v * 12.92 : 1.055 * std::pow(v, 1.0 / 2.4) - 0.055;
    }
    lut[1024] = lut[1023];
    for (int c = 0; c < 256; c++) {
        float v = (float)c / 0xff, s;
        if (v <= 0) s = 0;
        else if (v == 1) s = 1;
        else {
            float i = v * 1024;
            int64_t fl = f2i(i) & 1023;
            float v0 = (float)lut[fl], v1 = (float)lut[fl + 1];
            i -= (float)fl;
            s = v0 * (1.0f - i) + v1 * i;
        }
        out[c] = f2u8(s * 0xff + 0.5f);
    }
}
// The eight AO directions (material/ao.go:26-28), float32 angle accumulated.
void ao_dirs(float out[16]) {
    int k = 0;
    for (float a = 0.0f; a < (float)(2 * M_PI - 1e-4) && k < 8; a += (float)(M_PI / 4)) {
        out[2 * k] = (float)std::cos((double)a);
        out[2 * k + 1] = (float)std::sin((double)a);
        k++;
    }
}

// Tables of the AO march for frames up to 4096 x 4096 (k_ao):
//   [AO_TAB_OFF]  per direction, fl(dir.x * t) and fl(dir.y * t), t = 0..99;
//   [AO_TAB_DIST] per diagonal direction, the tap distance
//                 Len = sqrt32(fma64(ex, ex, ey * ey)) indexed by the binades
//                 (frexp exponents) of the tap coordinates and t.
// The second table rests on ex = X - fl(X + fl(dir.x * t)) being fixed by t
// and the binade of fl(X + fl(dir.x * t)) for every integer X in [0, 4096);
// that is checked here over every such X, and a failed check (or a binade
// outside the table) returns false, which keeps k_ao on its direct path.
uint32_t bitsof(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
bool ao_tables(const float dirs[16], std::vector<float>& tab) {
    tab.assign(AO_TAB_SIZE, __builtin_nanf(""));
    for (int d = 0; d < 8; d++)
        for (int t = 0; t < 100; t++) {
            tab[AO_TAB_OFF + d * 200 + t] = dirs[2 * d] * (float)t;
            tab[AO_TAB_OFF + d * 200 + 100 + t] = dirs[2 * d + 1] * (float)t;
        }
    double* rinv_t = reinterpret_cast<double*>(tab.data() + AO_TAB_T64);
    for (int t = 0; t < 100; t++) rinv_t[t] = 1.0 / (double)t;  // correctly rounded; t = 0 unused
    // axis directions: is the major offset exactly +-t, and from which minor
    // coordinate M on is fl(M + minor offset) == M for every t
    int32_t* axis = reinterpret_cast<int32_t*>(tab.data() + AO_TAB_AXIS);
    for (int a = 0; a < 4; a++) {
        const int d = 2 * a;
        const bool along_x = std::fabs(dirs[2 * d]) >= std::fabs(dirs[2 * d + 1]);
        const float maj = dirs[2 * d + (along_x ? 0 : 1)], mnr = dirs[2 * d + (along_x ? 1 : 0)];
        const int sgn = maj > 0 ? 1 : -1;
        bool exact = true;
        for (int t = 0; t < 100; t++) exact = exact && maj * (float)t == (float)(sgn * t);
        int mmin = 0;
        for (int M = 0; M < 4096; M++)
            for (int t = 0; t < 100; t++)
                if ((float)M + mnr * (float)t != (float)M) mmin = M + 1;
        axis[4 * a + 0] = along_x ? 1 : 0;
        axis[4 * a + 1] = sgn;
        axis[4 * a + 2] = exact ? mmin : 0x7fffffff;
        axis[4 * a + 3] = 0;
    }
    double* rinv_d = reinterpret_cast<double*>(tab.data() + AO_TAB_DIST);
    for (int q = 0; q < 4; q++) {
        const int d = 2 * q + 1;
        // e[comp][t][bucket]: the offset for each binade bucket (NaN: unused)
        std::vector<float> e(2 * 100 * AO_EXP_N, __builtin_nanf(""));
        for (int comp = 0; comp < 2; comp++)
            for (int t = 1; t < 100; t++) {
                const float c = dirs[2 * d + comp] * (float)t;
                for (int X = 0; X < 4096; X++) {
                    const float fx = (float)X, cx = fx + c, ex = fx - cx;
                    int k = 0;
                    (void)std::frexp(cx, &k);
                    const int b = k - AO_EXP_MIN;
                    if (b < 0 || b >= AO_EXP_N) return false;
                    float& slot = e[(comp * 100 + t) * AO_EXP_N + b];
                    if (slot != slot) slot = ex;
                    else if (bitsof(slot) != bitsof(ex)) return false;
                }
            }
        for (int bx = 0; bx < AO_EXP_N; bx++)
            for (int by = 0; by < AO_EXP_N; by++)
                for (int t = 1; t < 100; t++) {
                    const float ex = e[(0 * 100 + t) * AO_EXP_N + bx], ey = e[(1 * 100 + t) * AO_EXP_N + by];
                    if (ex != ex || ey != ey) continue;
                    const float s2 = (float)std::fma((double)ex, (double)ex, (double)(ey * ey));
                    const float len = std::sqrt(s2);
                    // from t = 2 on every distance is >= 1 (checked): the
                    // march tests distance < 1 at t = 1 only
                    if (t >= 2 && len < 1.0f) return false;
                    rinv_d[q * AO_DIST_PER_DIR + (bx * AO_EXP_N + by) * 100 + t] = 1.0 / (double)len;
                }
    }
    return true;
}

// createWeights8 (internal/imageutil/resize.go:148-164) with the linear
// kernel (:68-74) and filterLength 2, in float32 as the reference.
int resize_weights(int dy, float scale, std::vector<int16_t>& coeffs, std::vector<int32_t>& start) {
    const int fl = 2 * (int)max2((float)std::ceil((double)scale), 1.0f);
    const float ff = min2(1.0f / scale, 1.0f);
    coeffs.assign((size_t)dy * fl, 0);
    start.assign(dy, 0);
    for (int y = 0; y < dy; y++) {
        float interp = scale * ((float)y + 0.5f) - 0.5f;
        start[y] = (int)interp - fl / 2 + 1;
        interp -= (float)start[y];
        for (int i = 0; i < fl; i++) {
            float in = std::fabs((interp - (float)i) * ff);
            const float k = in <= 1 ? 1 - in : 0;
            coeffs[(size_t)y * fl + i] = (int16_t)(int32_t)(k * 256);
        }
    }
    return fl;
}

// calcFactors (internal/imageutil/resize.go:121-138) and the size Resize then
// makes (resize.go:17-23): a zero width or height follows the other's scale.
void resize_plan(int width, int height, int iw, int ih, float* sx, float* sy, int* ow, int* oh) {
    const float oldw = (float)iw, oldh = (float)ih;
    if (width == 0) {
        *sy = height == 0 ? 1.0f : oldh / (float)height;
        *sx = *sy;
    } else {
        *sx = oldw / (float)width;
        *sy = height == 0 ? *sx : oldh / (float)height;
    }
    if (width == 0) width = (int)(0.7f + oldw / *sx);
    if (height == 0) height = (int)(0.7f + oldh / *sy);
    *ow = width, *oh = height;
}

// math.Log2 as Go computes it (Frexp, then Log of the fraction)
double go_log2(double x) {
    if (!(x < INFINITY)) return x;
    int e;
    const double f = std::frexp(x, &e);
    if (f == 0.5) return (double)(e - 1);
    return std::log(f) * (1 / 0.693147180559945309417232121458176568) + (double)e;
}

template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(PRH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return PRH_OK;
}
template <class T>
int upload(T** p, const T* src, size_t n) {
    int rc = dalloc(p, n);
    if (rc) return rc;
    if (n) HIP_TRY(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return PRH_OK;
}

}  // namespace

struct prh_renderer;
// prh_render_batch spreads a batch over one stream per hardware queue the
// process has (GPU_MAX_HW_QUEUES, 4 unless set), at most PRH_BATCH_MAX.
constexpr int PRH_BATCH_MAX = 16;

struct prh_ctx {
    int device;
    hipStream_t stream;
    hipStream_t stream2;  // a frame's shadow passes, beside the main view's passes on `stream`
    int live_scenes;    // a context outlives every scene uploaded to it
    int live_surfaces;  // and every surface
    uint8_t* scratch;   // this device's `scratch` array (prh_diff_backward_noalloc), or null
    std::string scratch_err;  // why it is null
    // prh_render_batch: its streams and events (created on first use) and
    // the graph of the last captured batch, keyed by its renderers and their
    // work-buffer generations
    int nbs;  // batch streams in use
    hipStream_t bst[PRH_BATCH_MAX];
    hipEvent_t bev[2], bfork, bjoin[PRH_BATCH_MAX];
    bool batch_ready;
    hipGraph_t graph;
    hipGraphExec_t graph_exec;
    std::vector<std::pair<prh_renderer*, int>> graph_key;
    bool graph_wide;  // the graph replays wide launches
    // the last batch recorded for wide launches: its argument tables on the
    // device, keyed as the graph
    WideBatch* wide;
    uint8_t* wide_dev;
    std::vector<std::pair<prh_renderer*, int>> wide_key;
};

struct prh_scene {
    prh_ctx* ctx;
    int live_renderers;  // a scene outlives every renderer drawing it
    DevScene d;
    std::vector<void*> allocs;
    std::vector<M4> model;
    std::vector<prh_light> lights;
    std::vector<int> light_slot;
    int n_slots;
    float center[3];
    float ambient_sum;  // sum of the ambient intensities (differentiable shading)
    bool any_ao;        // some material has AmbientOcclusion
};

struct prh_renderer {
    prh_ctx* ctx;
    prh_scene* sc;
    prh_camera cam;
    prh_options opt;
    int W, H, tiles_x, tiles_y, n_tiles;  // the MSAA-sized framebuffer (width*MSAA x height*MSAA), shading tiles
    int rtile, rtiles_x, rtiles_y, n_rtiles;  // its raster tiles (prh_tuning.raster_tile)
    int msaa, OW, OH;                       // MSAA factor and the output image size
    int SW, SH, stiles_x, stiles_y, stiles_per_map, n_slots;  // shadow maps and their raster tiles
    ViewConsts vc;
    ShadowConsts shc;
    ShadeConsts sh;
    std::vector<void*> allocs;
    M4 *d_trans, *d_normal, *d_strans, *d_lview, *d_lproj;
    ShadowXf* d_lxf;
    float4 *tv_scr, *tv_nor, *tv_wp;
    float3* ts;  // [slot][vertex] light-space screen positions
    int4* tri_rec;     // per triangle: transformed vertices and flat material (once, k_tri_rec)
    uint32_t* tv_col;  // per transformed vertex: colour
    float2* tv_uv;     // and UV (scenes with textures)
    Prim* prims;
    uint2* pbox;  // pixel box per primitive slot
    ClipRec* clips;
    int32_t* tri_clip_slot;
    int32_t* blk_prims;  // per setup / clipper workgroup: unclipped primitives, pieces (statistics)
    int32_t* scan_sums;  // per scan block: primitive and work-item sums
    int32_t *tile_count, *tile_off, *chunk_off, *tile_cursor, *tile_list, *item_tile;
    ClipJob* clip_list;
    int32_t cap_items;
    // the shadow passes' own work set (they run beside the main view's)
    Prim* s_prims;
    uint2* s_pbox;
    int32_t *s_scan_sums, *s_tile_count, *s_tile_off, *s_chunk_off, *s_tile_cursor, *s_tile_list, *s_item_tile;
    int32_t s_cap_items;
    int32_t s_tile_slots;  // shadow slots x tiles per map
    hipEvent_t ev_fork, ev_join, ev_s[2];
    // split shading: resolve done, the shadow passes done, the AO's span
    hipEvent_t ev_res, ev_sdone, ev_sh[2];
    uint32_t* col_pre;            // split shading: the colour before AO
    uint8_t* post_flag;           // and what its second half does per pixel
    int32_t tile_slots;
    int32_t n_cu;
    uint32_t *first_all, *first_nan;
    uint64_t* vis;
    float* depth;
    uint32_t *image, *pregamma, *bg, *smaps;
    // MSAA > 1: k_shade writes image_big (W x H), Resize makes image (OW x OH)
    uint32_t *image_big, *resize_tmp;
    int16_t *rz_cx, *rz_cy;  // createWeights8 coefficients, both passes
    int32_t *rz_ox, *rz_oy;  // and their start offsets
    int rz_flx, rz_fly;
    float *linear, *bg_lin, *ao;
    int32_t *ao_xs, *ao_cnt;  // scenes with AO only: the rows' lists of AO pixels
    float* ao_tab;   // AO march tables (ao_tables)
    int32_t ao_fast; // the tables hold for this frame size
    Counters* cnt;      // [2]: main, shadow
    Counters* h_cnt;    // pinned
    Counters* d_h_cnt;  // its device address
    Capacities cap_main, cap_shadow;
    hipEvent_t ev[8];
    prh_frame_stats stats;
    bool has_frame;  // the visibility of a frame is in vis
    int work_gen;    // bumped when the work buffers are reallocated
    // differentiable shading, allocated on first use: parameters and
    // gradients in one block each (albedo, normal, kd, ks, shininess,
    // light_pos, light_int), the target, the forward output, the loss
    float *diff_par, *diff_grad, *diff_target, *diff_lin;
    double* diff_loss;
    hipEvent_t dev[2];
    uint8_t* diff_count;    // the fixed shadow visibility of frame count_serial
    uint32_t frame_serial;  // bumped by every completed frame
    uint32_t count_serial;
    int live_fits;  // a renderer outlives every prh_fit holding it as a view
    uint32_t* present_to;  // during prh_surface_render: the acquired surface texture
    prh_tuning tune;       // speed knobs (results do not depend on them)
    bool split_shade_last;  // the last frame split its shading
    // the last frame's per-pass times wait in its events until the stats are
    // read (prh_last_stats): a frame's own host work stays short
    bool times_pending, times_nan_rerun;
};

namespace {

void drop_graph(prh_ctx* ctx);

template <class T>
int ralloc(prh_renderer* r, T** p, size_t n) {
    int rc = dalloc(p, n);
    if (!rc) r->allocs.push_back((void*)*p);
    return rc;
}

void free_list(std::vector<void*>& v) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

// The raster tile grid of the main view and of the shadow maps for the
// renderer's raster_tile knob (8, 16 or 32), into its constants.
void raster_grid(prh_renderer* r) {
    const int T = r->tune.raster_tile, shift = T == 8 ? 3 : T == 32 ? 5 : 4;
    r->rtile = T;
    r->rtiles_x = (r->W + T - 1) / T, r->rtiles_y = (r->H + T - 1) / T;
    r->n_rtiles = r->rtiles_x * r->rtiles_y;
    r->stiles_x = (r->SW + T - 1) / T, r->stiles_y = (r->SH + T - 1) / T;
    r->stiles_per_map = r->stiles_x * r->stiles_y;
    r->vc.rtiles_x = r->rtiles_x, r->vc.rt_shift = shift;
    r->sh.v.rtiles_x = r->rtiles_x, r->sh.v.rt_shift = shift;
    r->shc.tiles_x = r->stiles_x, r->shc.tiles_y = r->stiles_y;
    r->shc.tiles_per_map = r->stiles_per_map, r->shc.rt_shift = shift;
}

int alloc_work(prh_renderer* r) {
    int rc;
    if ((rc = ralloc(r, &r->prims, (size_t)r->cap_main.prims))) return rc;
    if ((rc = ralloc(r, &r->pbox, (size_t)r->cap_main.prims))) return rc;
    if ((rc = ralloc(r, &r->clips, (size_t)r->cap_main.clip_subs))) return rc;
    if ((rc = ralloc(r, &r->tile_list, (size_t)r->cap_main.pairs))) return rc;
    if ((rc = ralloc(r, &r->clip_list, (size_t)r->cap_main.clip_tris))) return rc;
    // every item holds >= 1 primitive of one tile: items <= tiles + pairs / item_prims
    r->cap_items = r->tile_slots + r->cap_main.pairs / r->tune.item_prims + 1;
    if ((rc = ralloc(r, &r->item_tile, (size_t)r->cap_items))) return rc;
    if (r->n_slots > 0) {
        if ((rc = ralloc(r, &r->s_prims, (size_t)r->cap_shadow.prims))) return rc;
        if ((rc = ralloc(r, &r->s_pbox, (size_t)r->cap_shadow.prims))) return rc;
        if ((rc = ralloc(r, &r->s_tile_list, (size_t)r->cap_shadow.pairs))) return rc;
        r->s_cap_items = r->s_tile_slots + r->cap_shadow.pairs / r->tune.item_prims + 1;
        if ((rc = ralloc(r, &r->s_item_tile, (size_t)r->s_cap_items))) return rc;
    }
    return PRH_OK;
}
void free_work(prh_renderer* r) {
    for (void* p : {(void*)r->prims, (void*)r->pbox, (void*)r->clips, (void*)r->tile_list, (void*)r->clip_list,
                    (void*)r->item_tile, (void*)r->s_prims, (void*)r->s_pbox, (void*)r->s_tile_list,
                    (void*)r->s_item_tile}) {
        if (!p) continue;
        auto it = std::find(r->allocs.begin(), r->allocs.end(), p);
        if (it != r->allocs.end()) {
            (void)hipFree(p);
            r->allocs.erase(it);
        }
    }
}

}  // namespace

extern "C" {

const char* prh_last_error(void) { return g_err.c_str(); }

const char* prh_version(void) {
    static char v[96];
    snprintf(v, sizeof v, "libprh gfx950 (HIP %d.%d.%d)", HIP_VERSION_MAJOR, HIP_VERSION_MINOR, HIP_VERSION_PATCH);
    return v;
}

int prh_open(int32_t ordinal, prh_ctx** out) {
    if (!out) return fail(PRH_E_INVALID, "prh_open: null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PRH_E_NODEVICE, "no HIP device");
    if (ordinal < 0 || ordinal >= n) return fail(PRH_E_NODEVICE, "device ordinal out of range");
    HIP_TRY(hipSetDevice(ordinal));
    // (a failed lookup leaves the context usable: prh_diff_backward_noalloc
    // then reports it)
    uint8_t* scr = nullptr;
    std::string scr_err;
    if (scratch_setup(ordinal, &scr) != PRH_OK) scr = nullptr, scr_err = g_err;
    hipStream_t st, st2;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    prh_ctx* c = new prh_ctx();
    c->device = ordinal, c->stream = st, c->stream2 = st2, c->live_scenes = 0;
    c->scratch = scr, c->scratch_err = scr_err;
    *out = c;
    return PRH_OK;
}

int prh_close(prh_ctx* ctx) {
    if (!ctx) return fail(PRH_E_INVALID, "prh_close: null context");
    if (ctx->live_scenes > 0)
        return fail(PRH_E_INVALID, "prh_close: " + std::to_string(ctx->live_scenes) + " scene(s) still alive");
    if (ctx->live_surfaces > 0)
        return fail(PRH_E_INVALID, "prh_close: " + std::to_string(ctx->live_surfaces) + " surface(s) still alive");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    if (ctx->batch_ready) {
        for (int k = 0; k < ctx->nbs; k++) {
            (void)hipStreamSynchronize(ctx->bst[k]);
            (void)hipStreamDestroy(ctx->bst[k]);
            (void)hipEventDestroy(ctx->bjoin[k]);
        }
        (void)hipEventDestroy(ctx->bev[0]), (void)hipEventDestroy(ctx->bev[1]), (void)hipEventDestroy(ctx->bfork);
    }
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return PRH_OK;
}

// The textures and their mip chains (buffer.NewTexture, buffer/texture.go:
// 35-68): level 0 is the image, level i = imageutil.Resize(dx / 2^i, dy / 2^i)
// of it for int(Log2(max(dx, dy))) + 1 levels (one for a 1x1 image), each
// resize run on the device as two separable passes of k_resize.
static int upload_textures(prh_ctx* ctx, const prh_scene_desc* d, prh_scene* s, DevScene* ds) {
    if (d->n_textures == 0) return PRH_OK;
    struct Pass {
        int src_tex, level, iw, ih, w, h, flx, fly;
        size_t cx, ox, cy, oy;  // offsets into the coefficient / start tables
    };
    std::vector<DevTexture> tex(d->n_textures);
    std::vector<Pass> passes;
    std::vector<int16_t> coeffs;
    std::vector<int32_t> starts;
    int64_t total = 0, tmp_max = 1;
    for (int i = 0; i < d->n_textures; i++) {
        const prh_texture& t = d->textures[i];
        DevTexture& dt = tex[i];
        memset(&dt, 0, sizeof dt);
        dt.use_mipmap = t.use_mipmap != 0;
        dt.size = t.width;
        const int dx = t.width, dy = t.height;
        const int L = dx == 1 && dy == 1 ? 1 : (int)(float)go_log2((double)max2((float)dx, (float)dy)) + 1;
        if (L > TEX_MAX_LEVELS) return fail(PRH_E_UNSUPPORTED, "texture has too many mip levels");
        dt.n = L;
        dt.w[0] = dx, dt.h[0] = dy, dt.off[0] = (int32_t)total;
        total += (int64_t)dx * dy;
        for (int l = 1; l < L; l++) {
            float sx, sy;
            int w, h;
            resize_plan(dx >> l, dy >> l, dx, dy, &sx, &sy, &w, &h);
            dt.w[l] = w, dt.h[l] = h, dt.off[l] = (int32_t)total;
            total += (int64_t)w * h;
            if (total >= (1ll << 31)) return fail(PRH_E_UNSUPPORTED, "textures exceed 2^31 texels");
            Pass p{i, l, dx, dy, w, h, 0, 0, coeffs.size(), starts.size(), 0, 0};
            if (w == dx && h == dy) {  // Resize returns the image itself
                p.flx = 0;
                passes.push_back(p);
                continue;
            }
            std::vector<int16_t> cx, cy;
            std::vector<int32_t> ox, oy;
            p.flx = resize_weights(w, sx, cx, ox);
            p.fly = resize_weights(h, sy, cy, oy);
            // Go divides by the row's coefficient sum: it must not be zero
            for (auto [c, fl, n] : {std::tuple{&cx, p.flx, w}, std::tuple{&cy, p.fly, h}})
                for (int y = 0; y < n; y++) {
                    int32_t sum = 0;
                    for (int k = 0; k < fl; k++) sum += (*c)[(size_t)y * fl + k];
                    if (sum == 0) return fail(PRH_E_UNSUPPORTED, "texture resize with a zero filter sum");
                }
            coeffs.insert(coeffs.end(), cx.begin(), cx.end());
            starts.insert(starts.end(), ox.begin(), ox.end());
            p.cy = coeffs.size(), p.oy = starts.size();
            coeffs.insert(coeffs.end(), cy.begin(), cy.end());
            starts.insert(starts.end(), oy.begin(), oy.end());
            tmp_max = std::max<int64_t>(tmp_max, (int64_t)dy * w);
            passes.push_back(p);
        }
    }
    uint32_t* data = nullptr;
    DevTexture* dtex = nullptr;
    int rc = dalloc(&data, (size_t)total);
    if (!rc) s->allocs.push_back(data);
    if (!rc) rc = upload(&dtex, tex.data(), tex.size());
    if (!rc) s->allocs.push_back(dtex);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    for (int i = 0; i < d->n_textures && e == hipSuccess; i++)
        e = hipMemcpyAsync(data + tex[i].off[0], d->textures[i].rgba, (size_t)tex[i].w[0] * tex[i].h[0] * 4,
                           hipMemcpyHostToDevice, ctx->stream);
    int16_t* dco = nullptr;
    int32_t* dst = nullptr;
    uint32_t* tmp = nullptr;
    rc = upload(&dco, coeffs.data(), coeffs.size());
    if (!rc) rc = upload(&dst, starts.data(), starts.size());
    if (!rc) rc = dalloc(&tmp, (size_t)tmp_max);
    for (size_t k = 0; !rc && e == hipSuccess && k < passes.size(); k++) {
        const Pass& p = passes[k];
        const DevTexture& t = tex[p.src_tex];
        const uint32_t* img = data + t.off[0];
        uint32_t* dst_level = data + t.off[p.level];
        if (p.flx == 0) {
            e = hipMemcpyAsync(dst_level, img, (size_t)p.w * p.h * 4, hipMemcpyDeviceToDevice, ctx->stream);
            continue;
        }
        // rows of the image -> tmp (w rows of ih), then tmp's rows -> the level
        launch_resize(ctx->stream, img, p.iw, tmp, p.ih, p.w, dco + p.cx, dst + p.ox, p.flx);
        launch_resize(ctx->stream, tmp, p.ih, dst_level, p.w, p.h, dco + p.cy, dst + p.oy, p.fly);
    }
    if (!rc) {
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) rc = fail(PRH_E_HIP, std::string("texture mip levels: ") + hipGetErrorString(e));
    }
    (void)hipFree(dco), (void)hipFree(dst), (void)hipFree(tmp);
    if (rc) return rc;
    ds->n_textures = d->n_textures, ds->textures = dtex, ds->tex_data = data;
    return PRH_OK;
}

int prh_scene_upload(prh_ctx* ctx, const prh_scene_desc* d, prh_scene** out) {
    if (!ctx || !d || !out) return fail(PRH_E_INVALID, "prh_scene_upload: null argument");
    *out = nullptr;
    if (d->n_meshes <= 0 || d->n_objects <= 0) return fail(PRH_E_INVALID, "scene has no meshes or objects");
    if (d->n_lights < 0 || d->n_ambient < 0 || d->n_materials < 0) return fail(PRH_E_INVALID, "negative count");
    HIP_TRY(hipSetDevice(ctx->device));
    const int nm = d->n_meshes, O = d->n_objects;
    const int Vm = d->mesh_vtx_off[nm], Tm = d->mesh_tri_off[nm];
    for (int m = 0; m < nm; m++)
        if (d->mesh_vtx_off[m] > d->mesh_vtx_off[m + 1] || d->mesh_tri_off[m] > d->mesh_tri_off[m + 1])
            return fail(PRH_E_INVALID, "mesh offsets are not increasing");
    std::vector<float4> vpos(Vm), vnor(Vm);
    std::vector<uint32_t> vcol(Vm);
    for (int i = 0; i < Vm; i++) {
        vpos[i] = make_float4(d->vpos[4 * i], d->vpos[4 * i + 1], d->vpos[4 * i + 2], d->vpos[4 * i + 3]);
        vnor[i] = make_float4(d->vnor[4 * i], d->vnor[4 * i + 1], d->vnor[4 * i + 2], d->vnor[4 * i + 3]);
        memcpy(&vcol[i], d->vcol + 4 * i, 4);
    }
    std::vector<uint8_t> valid(Tm);
    for (int m = 0; m < nm; m++) {
        int nv = d->mesh_vtx_off[m + 1] - d->mesh_vtx_off[m];
        for (int t = d->mesh_tri_off[m]; t < d->mesh_tri_off[m + 1]; t++) {
            V4 p[3];
            for (int k = 0; k < 3; k++) {
                int lv = d->tri_v[3 * t + k];
                if (lv < 0 || lv >= nv) return fail(PRH_E_INVALID, "triangle vertex index out of range");
                const float* q = d->vpos + 4 * (d->mesh_vtx_off[m] + lv);
                p[k] = v4(q[0], q[1], q[2], q[3]);
            }
            valid[t] = tri_is_valid(p[0], p[1], p[2]) ? 1 : 0;
        }
    }
    std::vector<int32_t> tri_start(O + 1), vtx_start(O + 1);
    int64_t tg = 0, vt = 0;
    for (int o = 0; o < O; o++) {
        int m = d->obj_mesh[o];
        if (m < 0 || m >= nm) return fail(PRH_E_INVALID, "object mesh index out of range");
        tri_start[o] = (int32_t)tg, vtx_start[o] = (int32_t)vt;
        tg += d->mesh_tri_off[m + 1] - d->mesh_tri_off[m];
        vt += d->mesh_vtx_off[m + 1] - d->mesh_vtx_off[m];
    }
    if (tg >= (1ll << 29) || vt >= (1ll << 31)) return fail(PRH_E_UNSUPPORTED, "scene exceeds 2^29 triangles");
    tri_start[O] = (int32_t)tg, vtx_start[O] = (int32_t)vt;
    std::vector<int32_t> tri_obj(tg), vtx_obj(vt);
    for (int o = 0; o < O; o++) {
        for (int32_t g = tri_start[o]; g < tri_start[o + 1]; g++) tri_obj[g] = o;
        for (int32_t v = vtx_start[o]; v < vtx_start[o + 1]; v++) vtx_obj[v] = o;
    }
    if (d->n_textures < 0 || (d->n_textures > 0 && !d->textures)) return fail(PRH_E_INVALID, "bad texture list");
    for (int i = 0; i < d->n_textures; i++) {
        const prh_texture& t = d->textures[i];
        // NewTexture panics below 1x1 (buffer/texture.go:204-206); the bound
        // keeps the resize sums (<= 255 * 256 * scale) inside int32 as Go's
        if (t.width < 1 || t.height < 1 || !t.rgba) return fail(PRH_E_INVALID, "texture image is empty");
        if (t.width > 16384 || t.height > 16384) return fail(PRH_E_UNSUPPORTED, "texture larger than 16384");
    }
    if (d->n_textures > 0 && !d->vuv) return fail(PRH_E_INVALID, "textured scene without texture coordinates");
    std::vector<DevMaterial> mats(d->n_materials);
    bool any_ao = false;
    for (int i = 0; i < d->n_materials; i++) {
        const prh_material& m = d->materials[i];
        DevMaterial& dm = mats[i];
        for (int k = 0; k < 3; k++) {
            dm.diffuse[k] = (float)m.diffuse[k];
            dm.specular[k] = (float)m.specular[k];
            dm.tex[k] = (float)m.texture[k];
            float la = 0.0f;  // shader/blinn_cpu.go:48-52
            for (int e = 0; e < d->n_ambient; e++) la += d->ambient_intensity[e] * (float)m.texture[k];
            dm.la[k] = la;
        }
        dm.shininess = m.shininess;
        memcpy(&dm.tex_rgba, m.texture, 4);
        dm.flat_shading = m.flat_shading != 0;
        dm.ambient_occlusion = m.ambient_occlusion != 0;
        any_ao = any_ao || dm.ambient_occlusion;
        dm.receive_shadow = m.receive_shadow != 0;
        if (m.texture_id < -1 || m.texture_id >= d->n_textures)
            return fail(PRH_E_INVALID, "material texture_id out of range");
        dm.tex_id = m.texture_id;
    }
    prh_scene* s = new prh_scene();
    s->ctx = ctx;
    s->n_slots = 0;
    s->any_ao = any_ao;
    s->ambient_sum = 0.0f;
    for (int e = 0; e < d->n_ambient; e++) s->ambient_sum += d->ambient_intensity[e];
    std::vector<DevLight> lights(d->n_lights);
    for (int i = 0; i < d->n_lights; i++) {
        const prh_light& l = d->lights[i];
        if (l.kind != PRH_LIGHT_POINT && l.kind != PRH_LIGHT_DIRECTIONAL) {
            delete s;
            return fail(PRH_E_UNSUPPORTED, "light kind not supported");
        }
        DevLight& dl = lights[i];
        dl.kind = l.kind;
        for (int k = 0; k < 3; k++) {
            dl.pos[k] = l.position[k];
            dl.dir[k] = l.direction[k];
            dl.color[k] = (float)l.color[k];
        }
        dl.intensity = l.intensity;
        dl.cast_shadow = l.cast_shadow != 0;
        dl.slot = l.cast_shadow ? s->n_slots++ : -1;
        s->lights.push_back(l);
        s->light_slot.push_back(dl.slot);
    }
    for (int o = 0; o < O; o++) s->model.push_back(mat(d->obj_model + 16 * o));
    memcpy(s->center, d->center, sizeof s->center);

    DevScene& ds = s->d;
    memset(&ds, 0, sizeof ds);
    int rc = 0;
    auto up = [&](auto** p, const auto* src, size_t n) {
        if (rc) return;
        rc = upload(p, src, n);
        if (!rc) s->allocs.push_back((void*)*p);
    };
    float4 *dvpos, *dvnor;
    uint32_t* dvcol;
    int32_t *dtri_v, *dtri_mat, *dmvo, *dmto, *dom, *domb, *dts, *dvs, *dto, *dvo;
    uint8_t* dvalid;
    DevMaterial* dmats;
    DevLight* dlights;
    up(&dvpos, vpos.data(), Vm);
    up(&dvnor, vnor.data(), Vm);
    up(&dvcol, vcol.data(), Vm);
    up(&dtri_v, d->tri_v, (size_t)3 * Tm);
    up(&dtri_mat, d->tri_mat, Tm);
    up(&dvalid, valid.data(), Tm);
    up(&dmvo, d->mesh_vtx_off, nm + 1);
    up(&dmto, d->mesh_tri_off, nm + 1);
    up(&dom, d->obj_mesh, O);
    up(&domb, d->obj_mat_base, O);
    up(&dts, tri_start.data(), O + 1);
    up(&dvs, vtx_start.data(), O + 1);
    up(&dto, tri_obj.data(), tri_obj.size());
    up(&dvo, vtx_obj.data(), vtx_obj.size());
    up(&dmats, mats.data(), mats.size());
    up(&dlights, lights.data(), lights.size());
    float* damb;
    up(&damb, d->ambient_intensity, d->n_ambient);
    float2* dvuv = nullptr;
    if (d->n_textures > 0) {
        std::vector<float2> vuv(Vm);
        for (int i = 0; i < Vm; i++) vuv[i] = make_float2(d->vuv[2 * i], d->vuv[2 * i + 1]);
        up(&dvuv, vuv.data(), Vm);
    }
    if (!rc) rc = upload_textures(ctx, d, s, &ds);
    if (rc) {
        free_list(s->allocs);
        delete s;
        return rc;
    }
    ds.vpos = dvpos, ds.vnor = dvnor, ds.vcol = dvcol, ds.tri_v = dtri_v, ds.tri_mat = dtri_mat, ds.tri_valid = dvalid;
    ds.mesh_vtx_off = dmvo, ds.mesh_tri_off = dmto, ds.obj_mesh = dom, ds.obj_mat_base = domb;
    ds.obj_tri_start = dts, ds.obj_vtx_start = dvs, ds.tri_obj = dto, ds.vtx_obj = dvo;
    ds.n_objects = O, ds.n_tris = (int32_t)tg, ds.n_tverts = (int32_t)vt;
    ds.n_materials = d->n_materials, ds.mats = dmats;
    ds.n_lights = d->n_lights, ds.lights = dlights;
    ds.n_ambient = d->n_ambient, ds.ambient = damb, ds.vuv = dvuv;
    ctx->live_scenes++;
    *out = s;
    return PRH_OK;
}

int prh_scene_free(prh_scene* s) {
    if (!s) return fail(PRH_E_INVALID, "prh_scene_free: null scene");
    if (s->live_renderers > 0)
        return fail(PRH_E_INVALID, "prh_scene_free: " + std::to_string(s->live_renderers) + " renderer(s) still alive");
    s->ctx->live_scenes--;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_list(s->allocs);
    delete s;
    return PRH_OK;
}

int prh_renderer_new(prh_ctx* ctx, prh_scene* s, const prh_camera* cam, const prh_options* o, prh_renderer** out) {
    if (!ctx || !s || !cam || !o || !out) return fail(PRH_E_INVALID, "prh_renderer_new: null argument");
    *out = nullptr;
    if (s->ctx != ctx) return fail(PRH_E_INVALID, "prh_renderer_new: scene belongs to another context");
    const int msaa = o->msaa > 0 ? o->msaa : 1;
    if (o->width <= 0 || o->height <= 0 || msaa > 16 || (int64_t)o->width * msaa > 32768 ||
        (int64_t)o->height * msaa > 32768)
        return fail(PRH_E_INVALID, "framebuffer size out of range");
    if (o->shadow_size < 0 || o->shadow_size > 32768) return fail(PRH_E_INVALID, "shadow size out of range");
    if (cam->kind != PRH_CAMERA_PERSPECTIVE && cam->kind != PRH_CAMERA_ORTHOGRAPHIC)
        return fail(PRH_E_INVALID, "unknown camera kind");
    HIP_TRY(hipSetDevice(ctx->device));
    prh_renderer* r = new prh_renderer();
    r->ctx = ctx, r->sc = s, r->cam = *cam, r->opt = *o;
    r->tune = prh_tuning{8, ITEM_PRIMS, 64, 48, (size_t)o->width * o->height * msaa * msaa >= (1u << 19) ? 1 : 0,
                         TILE, 0.0f};
    // every pass renders the MSAA-sized buffer (render/raster.go:149-150)
    r->msaa = msaa, r->OW = o->width, r->OH = o->height;
    r->W = o->width * msaa, r->H = o->height * msaa;
    r->tiles_x = (r->W + TILE - 1) / TILE, r->tiles_y = (r->H + TILE - 1) / TILE;
    r->n_tiles = r->tiles_x * r->tiles_y;
    // shadow_size 0: framebuffer-sized maps as the reference (shadow.go:87-88)
    r->SW = o->shadow_size > 0 ? o->shadow_size : r->W;
    r->SH = o->shadow_size > 0 ? o->shadow_size : r->H;
    const bool shadows = o->shadow_map && s->n_slots > 0;
    r->n_slots = shadows ? s->n_slots : 0;

    // Main camera (render/raster.go:232-246,281-295).
    M4 view = view_matrix(v3a(cam->position), v3a(cam->target), v3a(cam->up));
    M4 proj = cam_proj(*cam);
    M4 vp = viewport_matrix((float)r->W, (float)r->H);
    M4 vinv = inv(view), pinv = inv(proj), vpinv = inv(vp);
    ViewConsts& vc = r->vc;
    vc.vp = vp, vc.vpinv = vpinv, vc.pinv = pinv, vc.vinv = vinv;
    vc.s2w = mulm(mulm(vinv, pinv), vpinv);
    memcpy(vc.cam, cam->position, sizeof vc.cam);
    vc.W = r->W, vc.H = r->H, vc.tiles_x = r->tiles_x, vc.tiles_y = r->tiles_y;
    vc.perspect = cam->kind == PRH_CAMERA_PERSPECTIVE;
    // the viewport box and clip planes of draw scale the (already MSAA-sized)
    // buffer by MSAA once more (render/raster.go:416-421,438-439)
    vc.vpbox_max_x = (float)(msaa * r->W), vc.vpbox_max_y = (float)(msaa * r->H);

    const int O = s->d.n_objects, L = s->d.n_lights;
    std::vector<M4> trans(O), normal(O);
    M4 pv = mulm(proj, view);
    for (int i = 0; i < O; i++) {
        trans[i] = mulm(pv, s->model[i]);          // raster.go:382
        normal[i] = transpose(inv(s->model[i]));   // raster.go:243
    }
    // Light cameras, render/shadow.go:33-90 (directional lights included).
    std::vector<M4> lview(std::max(1, L)), lproj(std::max(1, L)), strans((size_t)std::max(1, r->n_slots) * O);
    if (shadows) {
        V3 up = v3(0, 1, 0), center = v3a(s->center);
        for (int i = 0; i < L; i++) {
            if (!s->lights[i].cast_shadow) continue;
            V3 lp = v3a(s->lights[i].position);
            M4 lv = view_matrix(lp, center, up);
            M4 tm = mulm(mulm(lv, vinv), pinv);
            static const float corners[8][3] = {{1, 1, 1},  {1, 1, -1},  {1, -1, 1},  {-1, 1, 1},
                                                {-1, -1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};
            float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int k = 0; k < 8; k++) {
                V4 q = pos(apply(v4(corners[k][0], corners[k][1], corners[k][2], 1), tm));
                float qv[3] = {q.x, q.y, q.z};
                for (int a = 0; a < 3; a++) mn[a] = min2(mn[a], qv[a]), mx[a] = max2(mx[a], qv[a]);
            }
            lview[i] = lv;
            lproj[i] = ortho_proj(mn[0], mx[0], mn[1], mx[1], mx[2], mn[2] - 2);
            M4 lpv = mulm(lproj[i], lview[i]);
            int slot = s->light_slot[i];
            for (int o2 = 0; o2 < O; o2++) strans[(size_t)slot * O + o2] = mulm(lpv, s->model[o2]);
        }
    }
    ShadowConsts& shc = r->shc;
    shc.vp = viewport_matrix((float)r->SW, (float)r->SH);
    shc.SW = r->SW, shc.SH = r->SH;
    shc.n_slots = r->n_slots;
    shc.vpbox_max_x = (float)(msaa * r->SW), shc.vpbox_max_y = (float)(msaa * r->SH);
    shc.direct_max = r->tune.shadow_direct_max;

    raster_grid(r);
    ShadeConsts& sh = r->sh;
    memset(&sh, 0, sizeof sh);
    sh.v = vc;
    gamma_table(sh.gamma_lut);
    memcpy(&sh.background, o->background, 4);
    sh.shadow_map = o->shadow_map != 0;
    sh.gamma_correct = o->gamma_correct != 0;
    sh.n_lights = L;
    ao_dirs(sh.ao_dir);
    std::vector<float> aotab;
    const bool ao_tab_ok = ao_tables(sh.ao_dir, aotab);
    r->ao_fast = ao_tab_ok && r->W <= 4096 && r->H <= 4096;
    sh.SW = r->SW, sh.SH = r->SH;
    sh.svp = shc.vp;
    {
        // the zero pattern that lets the shading's shadow test skip zero
        // terms (checked here, entry by entry, for this renderer's matrices)
        auto last_row = [](const M4& m) { return m.a[12] == 0 && m.a[13] == 0 && m.a[14] == 0 && m.a[15] == 1; };
        auto diag_col = [&](const M4& m) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    if (i != j && m.a[4 * i + j] != 0) return false;
            return last_row(m);
        };
        bool sparse = diag_col(sh.svp);
        for (int i = 0; i < L && shadows; i++)
            if (s->lights[i].cast_shadow) sparse = sparse && last_row(lview[i]) && diag_col(lproj[i]);
        sh.shadow_sparse = sparse ? 1 : 0;
    }

    const size_t P = (size_t)r->W * r->H;
    const int Tg = s->d.n_tris, Vt = s->d.n_tverts;
    // primitive slots: one per triangle (per shadow slot), then the main
    // view's clipped fan pieces
    const int clip_tris = std::min(Tg, 1 << 16) + 16, clip0 = CLIP_PIECES * clip_tris;
    // the tile tables hold the smallest raster tile's grid, so the tile
    // knob never reallocates them
    const int rt_cap = ((r->W + RTILE_MIN - 1) / RTILE_MIN) * ((r->H + RTILE_MIN - 1) / RTILE_MIN);
    const int st_cap = ((r->SW + RTILE_MIN - 1) / RTILE_MIN) * ((r->SH + RTILE_MIN - 1) / RTILE_MIN);
    r->cap_main = Capacities{Tg + clip0, clip0, 4 * Tg + 4 * rt_cap + 1024, clip_tris};
    r->cap_shadow = Capacities{std::max(1, r->n_slots) * Tg, 0,
                               std::max(1, r->n_slots) * (2 * Tg + 4 * st_cap) + 1024, 0};
    r->tile_slots = rt_cap + 1;
    r->s_tile_slots = r->n_slots * st_cap + 1;
    int rc = 0;
    auto chk = [&](int v) {
        if (!rc) rc = v;
    };
    if (!rc) chk(upload(&r->d_trans, trans.data(), O));
    if (!rc) r->allocs.push_back(r->d_trans);
    if (!rc) chk(upload(&r->d_normal, normal.data(), O));
    if (!rc) r->allocs.push_back(r->d_normal);
    if (!rc) chk(upload(&r->d_strans, strans.data(), strans.size()));
    if (!rc) r->allocs.push_back(r->d_strans);
    if (!rc) chk(upload(&r->d_lview, lview.data(), lview.size()));
    if (!rc) r->allocs.push_back(r->d_lview);
    if (!rc) chk(upload(&r->d_lproj, lproj.data(), lproj.size()));
    if (!rc) r->allocs.push_back(r->d_lproj);
    {
        std::vector<ShadowXf> lxf(std::max(1, L));
        memset(lxf.data(), 0, lxf.size() * sizeof(ShadowXf));
        for (int i = 0; i < L; i++) {
            ShadowXf& x = lxf[i];
            for (int row = 0; row < 3; row++) {
                for (int col = 0; col < 3; col++) x.view[3 * row + col] = lview[i].a[4 * row + col];
                x.view_c[row] = lview[i].a[4 * row + 3];
                x.proj[row] = lproj[i].a[5 * row], x.proj_c[row] = lproj[i].a[4 * row + 3];
                x.svp[row] = sh.svp.a[5 * row], x.svp_c[row] = sh.svp.a[4 * row + 3];
            }
        }
        if (!rc) chk(upload(&r->d_lxf, lxf.data(), lxf.size()));
        if (!rc) r->allocs.push_back(r->d_lxf);
    }
    if (!rc) chk(upload(&r->ao_tab, aotab.data(), aotab.size()));
    if (!rc) r->allocs.push_back(r->ao_tab);
    if (!rc) chk(ralloc(r, &r->tv_scr, Vt));
    if (!rc) chk(ralloc(r, &r->tv_nor, Vt));
    if (!rc) chk(ralloc(r, &r->tv_wp, Vt));
    if (!rc) chk(ralloc(r, &r->tri_rec, Tg));
    if (!rc) chk(ralloc(r, &r->tv_col, Vt));
    if (!rc && s->d.n_textures > 0) chk(ralloc(r, &r->tv_uv, Vt));
    if (!rc) chk(ralloc(r, &r->ts, (size_t)std::max(1, r->n_slots) * Vt));
    if (!rc) chk(ralloc(r, &r->tri_clip_slot, Tg));
    if (!rc) chk(ralloc(r, &r->blk_prims, (size_t)setup_blocks(Tg) + CLIP_GRID));
    if (!rc) chk(ralloc(r, &r->scan_sums, 2 * (size_t)scan_blocks(r->tile_slots)));
    if (!rc) chk(ralloc(r, &r->tile_count, r->tile_slots));
    if (!rc) chk(ralloc(r, &r->tile_off, r->tile_slots + 1));
    if (!rc) chk(ralloc(r, &r->chunk_off, r->tile_slots + 1));

    if (!rc) chk(ralloc(r, &r->first_all, P));
    if (!rc) chk(ralloc(r, &r->first_nan, P));
    if (!rc) chk(ralloc(r, &r->tile_cursor, r->tile_slots));
    if (!rc && r->n_slots > 0) {
        chk(ralloc(r, &r->s_scan_sums, 2 * (size_t)scan_blocks(r->s_tile_slots)));
        if (!rc) chk(ralloc(r, &r->s_tile_count, r->s_tile_slots));
        if (!rc) chk(ralloc(r, &r->s_tile_off, r->s_tile_slots + 1));
        if (!rc) chk(ralloc(r, &r->s_chunk_off, r->s_tile_slots + 1));
        if (!rc) chk(ralloc(r, &r->s_tile_cursor, r->s_tile_slots));
    }
    if (!rc) chk(ralloc(r, &r->vis, P));
    if (!rc) chk(ralloc(r, &r->depth, P));
    if (!rc) chk(ralloc(r, &r->ao, P));
    if (!rc && s->any_ao) chk(ralloc(r, &r->ao_xs, P));
    if (!rc && s->any_ao) chk(ralloc(r, &r->ao_cnt, (size_t)r->H));
    if (!rc && s->any_ao && r->n_slots > 0) chk(ralloc(r, &r->col_pre, P));
    if (!rc && s->any_ao && r->n_slots > 0) chk(ralloc(r, &r->post_flag, P));
    if (!rc && msaa == 1) chk(ralloc(r, &r->image, P));
    if (!rc && msaa > 1) {
        // Resize(width, height, img): scale = old / new = MSAA (calcFactors,
        // resize.go:121-138)
        std::vector<int16_t> cx, cy;
        std::vector<int32_t> ox, oy;
        r->rz_flx = resize_weights(r->OW, (float)r->W / (float)r->OW, cx, ox);
        r->rz_fly = resize_weights(r->OH, (float)r->H / (float)r->OH, cy, oy);
        chk(ralloc(r, &r->image_big, P));
        if (!rc) chk(ralloc(r, &r->resize_tmp, (size_t)r->OW * r->H));
        if (!rc) chk(ralloc(r, &r->image, (size_t)r->OW * r->OH));
        if (!rc) chk(upload(&r->rz_cx, cx.data(), cx.size()));
        if (!rc) r->allocs.push_back(r->rz_cx);
        if (!rc) chk(upload(&r->rz_cy, cy.data(), cy.size()));
        if (!rc) r->allocs.push_back(r->rz_cy);
        if (!rc) chk(upload(&r->rz_ox, ox.data(), ox.size()));
        if (!rc) r->allocs.push_back(r->rz_ox);
        if (!rc) chk(upload(&r->rz_oy, oy.data(), oy.size()));
        if (!rc) r->allocs.push_back(r->rz_oy);
    }
    if (!rc && o->keep_linear) chk(ralloc(r, &r->pregamma, P));
    if (!rc && o->keep_linear) chk(ralloc(r, &r->linear, 3 * P));
    if (!rc) chk(ralloc(r, &r->bg, 4));
    if (!rc) chk(ralloc(r, &r->bg_lin, 4));
    if (!rc) chk(ralloc(r, &r->smaps, (size_t)std::max(1, r->n_slots) * r->SW * r->SH));
    if (!rc) chk(ralloc(r, &r->cnt, 2));
    if (!rc) chk(alloc_work(r));
    if (!rc) {
        hipError_t e = hipHostMalloc((void**)&r->h_cnt, 2 * sizeof(Counters), hipHostMallocDefault);
        if (e != hipSuccess) rc = fail(PRH_E_NOMEM, "hipHostMalloc failed");
        else if (hipHostGetDevicePointer((void**)&r->d_h_cnt, r->h_cnt, 0) != hipSuccess)
            rc = fail(PRH_E_HIP, "hipHostGetDevicePointer failed");
    }
    for (int i = 0; i < 8 && !rc; i++)
        if (hipEventCreate(&r->ev[i]) != hipSuccess) rc = fail(PRH_E_HIP, "hipEventCreate failed");
    for (int i = 0; i < 2 && !rc; i++)
        if (hipEventCreate(&r->ev_s[i]) != hipSuccess) rc = fail(PRH_E_HIP, "hipEventCreate failed");
    if (!rc && (hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&r->ev_res, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&r->ev_sdone, hipEventDisableTiming) != hipSuccess ||
                hipEventCreate(&r->ev_sh[0]) != hipSuccess || hipEventCreate(&r->ev_sh[1]) != hipSuccess))
        rc = fail(PRH_E_HIP, "hipEventCreate failed");
    if (rc) {
        free_list(r->allocs);
        delete r;
        return rc;
    }
    // Shadow maps start at zero once, here, as make([]float32) in initShadowMaps
    // (stored as ordered depth bits).
    launch_fill_u32(ctx->stream, r->smaps, (size_t)std::max(1, r->n_slots) * r->SW * r->SH, depth_ord(0.0f));
    launch_tri_rec(ctx->stream, s->d, r->tri_rec);
    // tile counts start at zero; every scan leaves them at zero
    launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(r->tile_count), r->tile_slots, 0u);
    if (r->n_slots > 0)
        launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(r->s_tile_count), r->s_tile_slots, 0u);
    {
        int n_cu = 0;
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || n_cu <= 0)
            n_cu = 256;
        r->n_cu = n_cu;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    memset(&r->stats, 0, sizeof r->stats);
    s->live_renderers++;
    *out = r;
    return PRH_OK;
}

int prh_renderer_free(prh_renderer* r) {
    if (!r) return fail(PRH_E_INVALID, "prh_renderer_free: null renderer");
    if (r->live_fits > 0)
        return fail(PRH_E_INVALID, "prh_renderer_free: " + std::to_string(r->live_fits) + " fit(s) still use it");
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    bool held = false;
    for (const auto& kv : r->ctx->graph_key) held |= kv.first == r;
    for (const auto& kv : r->ctx->wide_key) held |= kv.first == r;
    if (held) drop_graph(r->ctx);
    free_list(r->allocs);
    if (r->h_cnt) (void)hipHostFree(r->h_cnt);
    for (int i = 0; i < 8; i++)
        if (r->ev[i]) (void)hipEventDestroy(r->ev[i]);
    for (int i = 0; i < 2; i++) {
        if (r->dev[i]) (void)hipEventDestroy(r->dev[i]);
        if (r->ev_s[i]) (void)hipEventDestroy(r->ev_s[i]);
    }
    if (r->ev_fork) (void)hipEventDestroy(r->ev_fork);
    if (r->ev_join) (void)hipEventDestroy(r->ev_join);
    if (r->ev_res) (void)hipEventDestroy(r->ev_res);
    if (r->ev_sdone) (void)hipEventDestroy(r->ev_sdone);
    for (int i = 0; i < 2; i++)
        if (r->ev_sh[i]) (void)hipEventDestroy(r->ev_sh[i]);
    r->sc->live_renderers--;
    delete r;
    return PRH_OK;
}

}  // extern "C"

namespace {

int d_n_tris(const prh_renderer* r) { return r->sc->d.n_tris; }

// Where the frame's final image goes: the renderer's own, or the surface
// texture a prh_surface_render is drawing into.
uint32_t* final_image(const prh_renderer* r) { return r->present_to ? r->present_to : r->image; }

ShadeBuffers shade_buffers(const prh_renderer* r) {
    ShadeBuffers b;
    b.vis = r->vis, b.depth = r->depth, b.tv_scr = r->tv_scr, b.tv_nor = r->tv_nor, b.tv_wp = r->tv_wp;
    b.tri_rec = r->tri_rec, b.tv_col = r->tv_col, b.tv_uv = r->tv_uv;
    b.clips = r->clips, b.tri_clip_slot = r->tri_clip_slot, b.smaps = r->smaps;
    b.light_view = r->d_lview, b.light_proj = r->d_lproj, b.light_xf = r->d_lxf;
    b.image = r->msaa > 1 ? r->image_big : final_image(r), b.pregamma = r->pregamma, b.linear = r->linear;
    b.bg = r->bg, b.bg_lin = r->bg_lin;
    b.cnt = r->cnt;
    b.cnt_host = r->d_h_cnt;
    b.ao = r->ao;
    b.ao_xs = r->ao_xs, b.ao_cnt = r->ao_cnt;
    b.ao_tab = r->ao_tab, b.ao_fast = r->ao_fast;
    b.col_pre = nullptr, b.post_flag = nullptr;
    return b;
}

// Enqueues one frame; nan_mode re-runs the visibility raster in exact order.
// `timed` records the per-pass events (not while a batch is captured into a
// graph).  `wide`: the launches are being recorded into a wide batch (the
// stage numbers order its launches; the persistent rasters take fewer
// workgroups per view, since every view's run together).  A non-zero
// status means part of the frame may be enqueued and none of it is valid.
int enqueue_frame(prh_renderer* r, bool nan_mode, hipStream_t st, bool timed, bool wide = false) {
    const DevScene& d = r->sc->d;
    Counters* cm = r->cnt;
    Counters* cs = r->cnt + 1;
    const int persistent = wide ? std::max(32, r->n_cu * r->tune.raster_wg_per_cu / 16)
                                : r->n_cu * r->tune.raster_wg_per_cu;
    const int ip = r->tune.item_prims;
    if (timed) (void)hipEventRecord(r->ev[0], st);
    const size_t P = (size_t)r->W * r->H;
    // counters zeroed, visibility emptied (before the setup, which draws
    // small triangles into it, outside the exact-order pass, whose NaN rule
    // needs every fragment through the ordered raster): by the vertex passes
    const bool shadows = r->n_slots > 0 && !nan_mode;
    constexpr int32_t CNT_WORDS = (int32_t)(sizeof(Counters) / 4);
    // a plain frame runs its shadow passes on the context's second stream,
    // beside the main view's passes (they meet at the shading); a batch
    // keeps them on its own stream
    const bool split = shadows && !wide && st == r->ctx->stream && r->tune.overlap_shadows;
    const hipStream_t ss = split ? r->ctx->stream2 : st;
    // with AO as well, the AO runs on the second stream (after the shadow
    // passes) beside the shading's first half (everything but the AO factor)
    // on this one
    const bool split_shade = split && r->sc->any_ao && r->col_pre && r->post_flag;
    if (shadows) {
        const int nt = r->n_slots * r->stiles_per_map;
        if (split) {
            (void)hipEventRecord(r->ev_fork, st);
            (void)hipStreamWaitEvent(ss, r->ev_fork, 0);
        }
        if (timed) (void)hipEventRecord(r->ev_s[0], ss);
        wide_stage(2);
        launch_vertex_shadow(ss, d, r->d_strans, r->shc, r->ts, FrameClear{(uint32_t*)cs, CNT_WORDS, nullptr, 0});
        wide_stage(3);
        launch_setup_shadow(ss, d, r->shc, r->tri_rec, r->ts, r->s_prims, r->s_pbox, r->s_tile_count, cs, r->cap_shadow,
                            r->smaps);
        wide_stage(4);
        launch_scan(ss, r->s_tile_count, nt, r->s_tile_off, r->s_chunk_off, r->s_tile_cursor, r->s_scan_sums,
                    r->s_item_tile, r->s_cap_items, nullptr, 0, cs, ip);
        wide_stage(5);
        launch_bin(ss, r->s_pbox, cs, r->stiles_x, r->shc.rt_shift, r->stiles_per_map, 1, r->s_tile_cursor, r->s_tile_list,
                   r->cap_shadow, r->n_slots * d.n_tris, d.n_tris);
        RasterTarget rt{r->SW, r->SH, r->stiles_x, r->stiles_per_map, nt, nullptr, nullptr, nullptr, r->smaps};
        wide_stage(6);
        launch_raster(ss, 2, persistent, r->s_prims, r->s_tile_off, r->s_chunk_off, r->s_item_tile, r->s_cap_items,
                      r->s_tile_list, cs, rt, ip, r->rtile);
        if (timed) (void)hipEventRecord(r->ev_s[1], ss);
        if (split && !split_shade) (void)hipEventRecord(r->ev_join, ss);
    }
    if (timed) (void)hipEventRecord(r->ev[1], st);
    wide_stage(7);
    // (both counter blocks where no shadow pass runs)
    launch_vertex_main(st, d, r->d_trans, r->d_normal, r->vc, r->tv_scr, r->tv_nor, r->tv_wp, r->tv_col, r->tv_uv,
                       FrameClear{(uint32_t*)cm, (shadows ? 1 : 2) * CNT_WORDS, (uint64_t*)r->vis, P});
    if (timed) (void)hipEventRecord(r->ev[2], st);
    wide_stage(8);
    launch_setup_main(st, d, r->vc, r->tv_scr, r->tv_nor, r->prims, r->pbox, r->clips, r->tri_clip_slot, r->clip_list,
                      r->tile_count, r->blk_prims, cm, r->cap_main, (unsigned long long*)r->vis,
                      nan_mode ? 0 : r->tune.main_direct_max, r->tri_rec);
    wide_stage(9);
    launch_scan(st, r->tile_count, r->n_rtiles, r->tile_off, r->chunk_off, r->tile_cursor, r->scan_sums, r->item_tile,
                r->cap_items, r->blk_prims, d.n_tris ? setup_blocks(d.n_tris) + clip_grid(r->cap_main) : 0, cm, ip);
    wide_stage(10);
    launch_bin(st, r->pbox, cm, r->rtiles_x, r->vc.rt_shift, r->n_rtiles, 0, r->tile_cursor, r->tile_list, r->cap_main, d.n_tris,
               d.n_tris);
    if (timed) (void)hipEventRecord(r->ev[3], st);
    if (nan_mode) {
        launch_fill_u32(st, r->first_all, P, 0xffffffffu);
        launch_fill_u32(st, r->first_nan, P, 0xffffffffu);
    }
    RasterTarget rt{r->W, r->H, r->rtiles_x, r->n_rtiles, r->n_rtiles, (unsigned long long*)r->vis, r->first_all,
                    r->first_nan, nullptr};
    wide_stage(11);
    launch_raster(st, nan_mode ? 1 : 0, persistent, r->prims, r->tile_off, r->chunk_off, r->item_tile, r->cap_items,
                  r->tile_list, cm, rt, ip, r->rtile);
    wide_stage(12);
    launch_resolve(st, nan_mode, rt, cm, r->depth, d, r->sc->any_ao ? r->ao_xs : nullptr, r->ao_cnt, r->tri_rec);
    if (timed) (void)hipEventRecord(r->ev[4], st);
    const ShadeBuffers b = shade_buffers(r);
    // the split shading's buffers: both halves read and write col_pre
    ShadeBuffers bp = b;
    bp.col_pre = r->col_pre, bp.post_flag = r->post_flag;
    r->split_shade_last = split_shade;
    if (split_shade) {
        // (k_shade there finishes last when it runs on the second stream)
        (void)hipEventRecord(r->ev_sdone, ss);
        (void)hipEventRecord(r->ev_res, st);
        (void)hipStreamWaitEvent(ss, r->ev_res, 0);
        if (timed) (void)hipEventRecord(r->ev_sh[0], ss);
        launch_ao(ss, d, r->sh, b);
        if (timed) (void)hipEventRecord(r->ev_sh[1], ss);
        // the uncovered colour (pixel (0,0)'s shading, its AO factor
        // included) there too, beside the end of the shading's first half
        launch_shade_bg(ss, d, r->sh, b);
        (void)hipEventRecord(r->ev_join, ss);
        (void)hipStreamWaitEvent(st, r->ev_sdone, 0);  // the shading reads the shadow maps
        launch_shade(st, d, r->sh, bp);
    }
    wide_stage(13);
    if (r->sc->any_ao && !split_shade) launch_ao(st, d, r->sh, b);  // else no pixel reads an AO factor
    if (timed) (void)hipEventRecord(r->ev[5], st);
    if (split) (void)hipStreamWaitEvent(st, r->ev_join, 0);  // the shadow maps (and the AO factors)
    wide_stage(14);
    if (!split_shade) launch_shade_bg(st, d, r->sh, b);
    if (timed) (void)hipEventRecord(r->ev[7], st);
    wide_stage(15);
    if (!split_shade) launch_shade(st, d, r->sh, b);
    else if (!launch_shade_post(st, d, r->sh, bp))
        return fail(PRH_E_INVALID, "split shading without the AO factors or triangle records");
    if (r->msaa > 1) {
        // passAntialiasing (render/raster.go:361-378): Resize after gamma
        wide_stage(16);
        launch_resize(st, r->image_big, r->W, r->resize_tmp, r->H, r->OW, r->rz_cx, r->rz_ox, r->rz_flx);
        wide_stage(17);
        launch_resize(st, r->resize_tmp, r->H, final_image(r), r->OW, r->OH, r->rz_cy, r->rz_oy, r->rz_fly);
    }
    if (timed) (void)hipEventRecord(r->ev[6], st);
    // (k_shade_bg writes the counters into r->h_cnt)
    return PRH_OK;
}

// After a frame's work has completed: after an overflow, grows the work
// buffers and asks for another attempt (again = true); after a NaN depth,
// re-runs the visibility in exact order; then fills the statistics (per-pass
// times only if the frame recorded its events).
int finish_frame(prh_renderer* r, int attempt, bool& again, bool timed = true) {
    again = false;
    hipStream_t st = r->ctx->stream;
    const Counters cm = r->h_cnt[0], cs = r->h_cnt[1];
    bool grow = false;
    if (cs.overflow || cs.n_pairs > r->cap_shadow.pairs) {
        r->cap_shadow.pairs = std::max(r->cap_shadow.pairs, cs.n_pairs + cs.n_pairs / 4 + 16);
        grow = true;
    }
    if (cm.overflow || cm.n_pairs > r->cap_main.pairs || cm.n_clipped > r->cap_main.clip_tris) {
        r->cap_main.pairs = std::max(r->cap_main.pairs, cm.n_pairs + cm.n_pairs / 4 + 16);
        r->cap_main.clip_tris = std::max(r->cap_main.clip_tris, cm.n_clipped + cm.n_clipped / 4 + 16);
        r->cap_main.clip_subs = std::max(r->cap_main.clip_subs, CLIP_PIECES * r->cap_main.clip_tris);
        r->cap_main.prims = d_n_tris(r) + r->cap_main.clip_subs;
        grow = true;
    }
    if (grow) {
        if (attempt >= 3) return fail(PRH_E_NOMEM, "work buffers keep overflowing");
        free_work(r);
        int rc = alloc_work(r);
        if (rc) return rc;
        r->work_gen++;
        again = true;
        return PRH_OK;
    }
    if (cs.shadow_nan)
        return fail(PRH_E_UNSUPPORTED, "NaN depth in a shadow pass: its result depends on sequential order");
    bool nan_rerun = false;
    if (cm.nan_flag) {
        const int rc = enqueue_frame(r, true, st, timed);
        HIP_TRY(hipStreamSynchronize(st));
        if (rc) return rc;
        nan_rerun = true;
    }
    prh_frame_stats& S = r->stats;
    S.ms_total = S.ms_shadow = S.ms_vertex = S.ms_setup = S.ms_raster = S.ms_ao = S.ms_shade = 0.0f;
    r->times_pending = timed, r->times_nan_rerun = nan_rerun;
    S.n_prims = cm.n_prims, S.n_clipped = cm.n_clipped, S.n_bin_pairs = cm.n_pairs;
    S.nan_rerun = nan_rerun ? 1 : 0;
    r->has_frame = true;
    r->frame_serial++;
    return PRH_OK;
}

void drop_graph_exec(prh_ctx* ctx) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr, ctx->graph = nullptr;
    ctx->graph_key.clear();
}

// The cached graph and the wide tables (a wide graph reads them).
void drop_graph(prh_ctx* ctx) {
    drop_graph_exec(ctx);
    if (ctx->wide) wide_free(ctx->wide);
    if (ctx->wide_dev) {
        (void)hipStreamSynchronize(ctx->stream);
        for (int k = 0; k < ctx->nbs; k++) (void)hipStreamSynchronize(ctx->bst[k]);
        (void)hipFree(ctx->wide_dev);
    }
    ctx->wide = nullptr, ctx->wide_dev = nullptr;
    ctx->wide_key.clear();
}

// Stream of each view of a batch: largest estimated cost first, each to the
// least loaded stream (the cost: pixels plus triangles for every pass).
std::vector<int> batch_lanes(prh_renderer* const* rs, int n) {
    std::vector<int> order(n), lane(n);
    std::vector<double> cost(n), load(rs[0]->ctx->nbs, 0.0);
    for (int i = 0; i < n; i++) {
        order[i] = i;
        cost[i] = (double)rs[i]->W * rs[i]->H + (double)rs[i]->sc->d.n_tris * (1 + rs[i]->n_slots) * 4.0 +
                  (double)rs[i]->n_slots * rs[i]->SW * rs[i]->SH * 0.25;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lane[i] = k;
        load[k] += cost[i];
    }
    return lane;
}

int batch_streams(prh_ctx* ctx) {
    if (ctx->batch_ready) return PRH_OK;
    const char* q = getenv("GPU_MAX_HW_QUEUES");
    ctx->nbs = std::min(PRH_BATCH_MAX, std::max(1, q ? atoi(q) : 4));
    for (int k = 0; k < ctx->nbs; k++) {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->bst[k], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->bjoin[k], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreate(&ctx->bev[0]));
    HIP_TRY(hipEventCreate(&ctx->bev[1]));
    HIP_TRY(hipEventCreateWithFlags(&ctx->bfork, hipEventDisableTiming));
    ctx->batch_ready = true;
    return PRH_OK;
}

// Records one frame of every renderer as wide launches and uploads their
// argument tables (kept until the batch, or a renderer's work buffers,
// change).
int batch_wide(prh_ctx* ctx, prh_renderer* const* rs, int n) {
    std::vector<std::pair<prh_renderer*, int>> key(n);
    for (int i = 0; i < n; i++) key[i] = {rs[i], rs[i]->work_gen};
    if (ctx->wide && key == ctx->wide_key) return PRH_OK;
    drop_graph(ctx);
    WideBatch* w = wide_new();
    wide_record(w);
    int rc = PRH_OK;
    for (int i = 0; i < n && !rc; i++) rc = enqueue_frame(rs[i], false, ctx->bst[0], false, true);
    wide_record(nullptr);
    if (rc) {
        wide_free(w);
        return rc;
    }
    const size_t bytes = wide_layout(w);
    if (!bytes) {
        wide_free(w);
        return fail(PRH_E_HIP, "wide batch: a stage mixes kernels");
    }
    uint8_t* dev = nullptr;
    if (hipMalloc((void**)&dev, bytes) != hipSuccess) {
        wide_free(w);
        return fail(PRH_E_NOMEM, "wide batch: hipMalloc failed");
    }
    if (wide_upload(w, dev)) {
        (void)hipFree(dev);
        wide_free(w);
        return fail(PRH_E_HIP, "wide batch: upload failed");
    }
    ctx->wide = w, ctx->wide_dev = dev, ctx->wide_key = key;
    return PRH_OK;
}

// Captures one frame of every renderer into a graph (kept until the batch,
// a renderer's work buffers or the mode change): the wide launches, or each
// view's own launches spread over the batch streams.
int batch_graph(prh_ctx* ctx, prh_renderer* const* rs, int n, bool wide) {
    std::vector<std::pair<prh_renderer*, int>> key(n);
    for (int i = 0; i < n; i++) key[i] = {rs[i], rs[i]->work_gen};
    if (ctx->graph_exec && key == ctx->graph_key && ctx->graph_wide == wide) return PRH_OK;
    drop_graph_exec(ctx);
    if (wide) {
        int rc = batch_wide(ctx, rs, n);
        if (rc) return rc;
    }
    hipStream_t s0 = ctx->bst[0];
    HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    int rc = PRH_OK;
    if (wide) {
        wide_launch(ctx->wide, s0, ctx->wide_dev);
    } else {
        e = hipEventRecord(ctx->bfork, s0);
        for (int k = 1; k < ctx->nbs && e == hipSuccess; k++) e = hipStreamWaitEvent(ctx->bst[k], ctx->bfork, 0);
        if (e == hipSuccess) {
            const std::vector<int> lane = batch_lanes(rs, n);
            for (int i = 0; i < n && !rc; i++) rc = enqueue_frame(rs[i], false, ctx->bst[lane[i]], false);
            for (int k = 1; k < ctx->nbs && e == hipSuccess; k++) {
                e = hipEventRecord(ctx->bjoin[k], ctx->bst[k]);
                if (e == hipSuccess) e = hipStreamWaitEvent(s0, ctx->bjoin[k], 0);
            }
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s0, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(PRH_E_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        ctx->graph_exec = nullptr;
        return fail(PRH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    ctx->graph = g;
    ctx->graph_key = key;
    ctx->graph_wide = wide;
    return PRH_OK;
}

// Offsets (in floats) of the differentiable parameters / gradients block.
struct DiffLayout {
    size_t albedo, normal, kd, ks, shin, lpos, lint, total;
};
DiffLayout diff_layout(const prh_renderer* r) {
    const size_t V = r->sc->d.n_tverts, M = r->sc->d.n_materials, L = r->sc->d.n_lights;
    DiffLayout l;
    l.albedo = 0, l.normal = 3 * V, l.kd = 6 * V, l.ks = l.kd + 3 * M, l.shin = l.ks + 3 * M;
    l.lpos = l.shin + M, l.lint = l.lpos + 3 * L, l.total = l.lint + L + 1;
    return l;
}

// The fixed shadow visibility of the renderer's last frame (k_shadow_count),
// computed once per frame; null when the frame draws no shadow maps.
int shadow_counts(prh_renderer* r, hipStream_t st, const uint8_t** out) {
    *out = nullptr;
    if (!r->sh.shadow_map) return PRH_OK;
    if (!r->diff_count) {
        int rc = ralloc(r, &r->diff_count, (size_t)r->W * r->H);
        if (rc) return rc;
        r->count_serial = r->frame_serial - 1;
    }
    if (r->count_serial != r->frame_serial) {
        launch_shadow_count(st, r->sc->d, r->sh, shade_buffers(r), r->diff_count);
        HIP_TRY(hipGetLastError());
        r->count_serial = r->frame_serial;
    }
    *out = r->diff_count;
    return PRH_OK;
}

// A view's differentiable-shading buffers over a parameter block `par` and a
// gradient block `grad` (both laid out by diff_layout).
DiffBuffers diff_buffers(const prh_renderer* r, const DiffLayout& l, const float* par, float* grad, double* loss,
                         bool normal, const uint8_t* count) {
    DiffBuffers b;
    b.vis = r->vis, b.tv_scr = r->tv_scr, b.tv_nor = r->tv_nor, b.tv_wp = r->tv_wp;
    b.albedo = par + l.albedo, b.normal = normal ? par + l.normal : nullptr;
    b.kd = par + l.kd, b.ks = par + l.ks, b.shininess = par + l.shin;
    b.light_pos = par + l.lpos, b.light_int = par + l.lint;
    b.ambient = r->sc->ambient_sum;
    b.shadow_count = count, b.gbuf = nullptr;
    b.target = nullptr, b.linear = nullptr;
    b.g_albedo = grad + l.albedo, b.g_normal = grad + l.normal, b.g_kd = grad + l.kd;
    b.g_ks = grad + l.ks, b.g_shin = grad + l.shin, b.g_lpos = grad + l.lpos, b.g_lint = grad + l.lint;
    b.loss = loss;
    b.want = DIFF_ALL;
    return b;
}

// Uploads the parameter arrays of `p` into a block laid out by diff_layout.
int upload_params(const prh_renderer* r, const DiffLayout& l, const prh_diff_params* p, float* par, hipStream_t st) {
    const DevScene& d = r->sc->d;
    const size_t V = d.n_tverts, M = d.n_materials, L = d.n_lights;
    auto up = [&](size_t off, const float* src, size_t n) -> hipError_t {
        return n ? hipMemcpyAsync(par + off, src, n * 4, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    HIP_TRY(up(l.albedo, p->albedo, 3 * V));
    if (p->normal) HIP_TRY(up(l.normal, p->normal, 3 * V));
    HIP_TRY(up(l.kd, p->kd, 3 * M));
    HIP_TRY(up(l.ks, p->ks, 3 * M));
    HIP_TRY(up(l.shin, p->shininess, M));
    HIP_TRY(up(l.lpos, p->light_pos, 3 * L));
    HIP_TRY(up(l.lint, p->light_int, L));
    return PRH_OK;
}

int check_params(const prh_renderer* r, const prh_diff_params* p) {
    const DevScene& d = r->sc->d;
    if (d.n_lights > DIFF_MAX_LIGHTS) return fail(PRH_E_UNSUPPORTED, "differentiable shading takes at most 8 lights");
    if ((d.n_tverts && !p->albedo) || (d.n_materials && (!p->kd || !p->ks || !p->shininess)) ||
        (d.n_lights && (!p->light_pos || !p->light_int)))
        return fail(PRH_E_INVALID, "prh_diff_params: missing parameter array");
    return PRH_OK;
}

// One launch of the differentiable shading: forward (target == nullptr,
// result in diff_lin) or adjoint (loss and gradients copied to the host).
int diff_run(prh_renderer* r, const prh_diff_params* p, const float* target, double* loss,
             const prh_diff_grads* grads) {
    const DevScene& d = r->sc->d;
    if (!r->has_frame) return fail(PRH_E_INVALID, "differentiable shading needs a rendered frame (prh_render)");
    int rc0 = check_params(r, p);
    if (rc0) return rc0;
    hipStream_t st = r->ctx->stream;
    const DiffLayout l = diff_layout(r);
    const size_t P = (size_t)r->W * r->H;
    if (!r->diff_par) {
        int rc = 0;
        auto chk = [&](int v) {
            if (!rc) rc = v;
        };
        chk(ralloc(r, &r->diff_par, l.total));
        if (!rc) chk(ralloc(r, &r->diff_grad, l.total));
        if (!rc) chk(ralloc(r, &r->diff_target, 3 * P));
        if (!rc) chk(ralloc(r, &r->diff_lin, 3 * P));
        if (!rc) chk(ralloc(r, &r->diff_loss, 1));
        for (int i = 0; i < 2 && !rc; i++)
            if (hipEventCreate(&r->dev[i]) != hipSuccess) rc = fail(PRH_E_HIP, "hipEventCreate failed");
        if (rc) {
            r->diff_par = nullptr;  // (the allocations stay on the renderer's list)
            return rc;
        }
    }
    const size_t V = d.n_tverts, M = d.n_materials, L = d.n_lights;
    int rc1 = upload_params(r, l, p, r->diff_par, st);
    const uint8_t* count = nullptr;
    if (!rc1) rc1 = shadow_counts(r, st, &count);
    if (rc1) return rc1;
    DiffBuffers b = diff_buffers(r, l, r->diff_par, r->diff_grad, r->diff_loss, p->normal != nullptr, count);
    b.target = r->diff_target, b.linear = r->diff_lin;
    const bool backward = target != nullptr;
    if (backward) {
        HIP_TRY(hipMemcpyAsync(r->diff_target, target, 12 * P, hipMemcpyHostToDevice, st));
        launch_fill_u32(st, reinterpret_cast<uint32_t*>(r->diff_grad), l.total, 0u);
        launch_fill_u64(st, reinterpret_cast<uint64_t*>(r->diff_loss), 1, 0ull);
    }
    (void)hipEventRecord(r->dev[0], st);
    launch_diff(st, backward, d, r->vc, b);
    (void)hipEventRecord(r->dev[1], st);
    if (backward) {
        auto down = [&](float* dst, size_t off, size_t n) -> hipError_t {
            return dst && n ? hipMemcpyAsync(dst, r->diff_grad + off, n * 4, hipMemcpyDeviceToHost, st) : hipSuccess;
        };
        HIP_TRY(down(grads->albedo, l.albedo, 3 * V));
        HIP_TRY(down(grads->normal, l.normal, 3 * V));
        HIP_TRY(down(grads->kd, l.kd, 3 * M));
        HIP_TRY(down(grads->ks, l.ks, 3 * M));
        HIP_TRY(down(grads->shininess, l.shin, M));
        HIP_TRY(down(grads->light_pos, l.lpos, 3 * L));
        HIP_TRY(down(grads->light_int, l.lint, L));
        HIP_TRY(hipMemcpyAsync(loss, r->diff_loss, 8, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    (void)hipEventElapsedTime(&r->stats.ms_diff, r->dev[0], r->dev[1]);
    return PRH_OK;
}

}  // namespace

// ---- headless present: a ring of RGBA8 textures in HBM
struct prh_surface {
    prh_ctx* ctx;
    int32_t w, h, frames;
    uint32_t* tex[16];
    int64_t frame;  // frames presented
    bool acquired;
};

namespace {

void surface_release(prh_surface* s) {
    for (int k = 0; k < s->frames; k++)
        if (s->tex[k]) (void)hipFree(s->tex[k]), s->tex[k] = nullptr;
}

int surface_alloc(prh_surface* s, int32_t w, int32_t h) {
    s->w = w, s->h = h;
    for (int k = 0; k < s->frames; k++) {
        int rc = dalloc(&s->tex[k], (size_t)w * h);
        if (rc) {
            surface_release(s);
            return rc;
        }
    }
    return PRH_OK;
}

}  // namespace

// ---- batched inverse rendering (prh_fit_*)
struct prh_fit {
    prh_ctx* ctx;
    prh_scene* sc;
    std::vector<prh_renderer*> views;
    std::vector<uint32_t> serials;  // each view's frame at creation
    bool stored;
    DiffLayout l;
    std::vector<void*> allocs;
    float *par, *grad, *base, *theta, *m, *v, *lower;
    int32_t* idx;
    double *loss, *hist;
    int32_t hist_cap;
    int32_t* step;
    std::vector<float*> target;
    std::vector<const uint8_t*> count;
    std::vector<float4*> gbuf;
    DiffView* d_views;
    int32_t total_blocks;
    hipEvent_t ev[4];
    prh_fit_stats stats;
};

namespace {

template <class T>
int falloc(prh_fit* f, T** p, size_t n) {
    int rc = dalloc(p, n);
    if (!rc) f->allocs.push_back((void*)*p);
    return rc;
}

int fit_check(prh_fit* f) {
    for (size_t i = 0; i < f->views.size(); i++)
        if (!f->views[i]->has_frame || f->views[i]->frame_serial != f->serials[i])
            return fail(PRH_E_INVALID, "prh_fit: a view was rendered again after prh_fit_new");
    return PRH_OK;
}

// Every view's buffers, launch blocks in view order, uploaded to d_views.
int fit_views(prh_fit* f, bool backward, uint32_t want, hipStream_t st) {
    std::vector<DiffView> h(f->views.size());
    int32_t blk = 0;
    for (size_t i = 0; i < f->views.size(); i++) {
        prh_renderer* r = f->views[i];
        DiffBuffers b = diff_buffers(r, f->l, f->par, f->grad, f->loss, false, f->count[i]);
        if (backward) {
            b.target = f->target[i];
            b.gbuf = f->stored ? f->gbuf[i] : nullptr;
        } else {
            b.linear = f->target[i];
        }
        b.want = want;
        h[i].v = r->vc, h[i].b = b;
        h[i].block0 = blk, h[i].nblocks = diff_blocks(r->vc);
        blk += h[i].nblocks;
    }
    f->total_blocks = blk;
    HIP_TRY(hipMemcpyAsync(f->d_views, h.data(), h.size() * sizeof(DiffView), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // h is a local
    return PRH_OK;
}

// The optimised indices of the parameter block and their lower bounds.
void fit_groups(const prh_fit* f, int32_t groups, std::vector<int32_t>& idx, std::vector<float>& lower) {
    const DevScene& d = f->sc->d;
    const size_t V = d.n_tverts, M = d.n_materials, L = d.n_lights;
    const float inf = std::numeric_limits<float>::infinity();
    auto add = [&](int32_t bit, size_t off, size_t n, float lo) {
        if (!(groups & bit)) return;
        for (size_t k = 0; k < n; k++) idx.push_back((int32_t)(off + k)), lower.push_back(lo);
    };
    add(PRH_FIT_ALBEDO, f->l.albedo, 3 * V, 0.0f);
    add(PRH_FIT_KD, f->l.kd, 3 * M, 0.0f);
    add(PRH_FIT_KS, f->l.ks, 3 * M, 0.0f);
    add(PRH_FIT_SHININESS, f->l.shin, M, 0.0f);
    add(PRH_FIT_LIGHT_POS, f->l.lpos, 3 * L, -inf);
    add(PRH_FIT_LIGHT_INT, f->l.lint, L, 0.0f);
}

}  // namespace

extern "C" {

int prh_render(prh_renderer* r) {
    if (!r) return fail(PRH_E_INVALID, "prh_render: null renderer");
    r->has_frame = false;
    HIP_TRY(hipSetDevice(r->ctx->device));
    hipStream_t st = r->ctx->stream;
    for (int attempt = 0;; attempt++) {
        const int rc0 = enqueue_frame(r, false, st, true);
        HIP_TRY(hipStreamSynchronize(st));
        if (rc0) return rc0;
        HIP_TRY(hipGetLastError());
        bool again = false;
        int rc = finish_frame(r, attempt, again);
        if (rc || !again) return rc;
    }
}

int prh_render_batch(prh_renderer* const* rs, int32_t n, int32_t flags, float* ms) {
    if (!rs || n < 0) return fail(PRH_E_INVALID, "prh_render_batch: null renderer list");
    if (n == 0) return PRH_OK;
    prh_ctx* ctx = rs[0] ? rs[0]->ctx : nullptr;
    for (int i = 0; i < n; i++)
        if (!rs[i] || rs[i]->ctx != ctx) return fail(PRH_E_INVALID, "prh_render_batch: renderers of one context");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = batch_streams(ctx);
    if (rc) return rc;
    if (flags & ~(PRH_BATCH_GRAPH | PRH_BATCH_STREAMS)) return fail(PRH_E_INVALID, "prh_render_batch: unknown flags");
    const bool graph = (flags & PRH_BATCH_GRAPH) != 0;
    const bool wide = (flags & PRH_BATCH_STREAMS) == 0;
    hipStream_t s0 = ctx->bst[0];
    if (graph) rc = batch_graph(ctx, rs, n, wide);
    else if (wide) rc = batch_wide(ctx, rs, n);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ctx->bev[0], s0));
    if (graph) {
        HIP_TRY(hipGraphLaunch(ctx->graph_exec, s0));
    } else if (wide) {
        wide_launch(ctx->wide, s0, ctx->wide_dev);
        HIP_TRY(hipGetLastError());
    } else {
        // fork: every batch stream starts after the work already on s0
        for (int k = 1; k < ctx->nbs; k++) HIP_TRY(hipStreamWaitEvent(ctx->bst[k], ctx->bev[0], 0));
        const std::vector<int> lane = batch_lanes(rs, n);
        for (int i = 0; i < n && !rc; i++) {
            rs[i]->has_frame = false;
            rc = enqueue_frame(rs[i], false, ctx->bst[lane[i]], true);
        }
        for (int k = 1; k < ctx->nbs; k++) {
            HIP_TRY(hipEventRecord(ctx->bjoin[k], ctx->bst[k]));
            HIP_TRY(hipStreamWaitEvent(s0, ctx->bjoin[k], 0));
        }
    }
    HIP_TRY(hipEventRecord(ctx->bev[1], s0));
    HIP_TRY(hipEventSynchronize(ctx->bev[1]));
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    if (ms) (void)hipEventElapsedTime(ms, ctx->bev[0], ctx->bev[1]);
    // per view: the same checks as prh_render; a view that overflowed its
    // work buffers (or met a NaN depth) is rendered again on its own
    for (int i = 0; i < n; i++) {
        prh_renderer* r = rs[i];
        if (graph || wide) r->has_frame = false;
        bool again = false;
        rc = finish_frame(r, 0, again, !graph && !wide);
        if (rc) return rc;
        if (again && (rc = prh_render(r))) return rc;
    }
    return PRH_OK;
}

int prh_read_image(prh_renderer* r, uint8_t* rgba, size_t bytes) {
    if (!r || !rgba) return fail(PRH_E_INVALID, "prh_read_image: null argument");
    size_t need = (size_t)4 * r->OW * r->OH;
    if (bytes < need) return fail(PRH_E_INVALID, "prh_read_image: buffer too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(hipMemcpyAsync(rgba, r->image, need, hipMemcpyDeviceToHost, r->ctx->stream));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    return PRH_OK;
}

int prh_read_buffer(prh_renderer* r, int32_t which, void* dst, size_t bytes) {
    if (!r || !dst) return fail(PRH_E_INVALID, "prh_read_buffer: null argument");
    const size_t P = (size_t)r->W * r->H;
    const void* src = nullptr;
    size_t need = 0;
    switch (which) {
        case PRH_BUF_VIS: src = r->vis, need = 8 * P; break;
        case PRH_BUF_DEPTH: src = r->depth, need = 4 * P; break;
        case PRH_BUF_LINEAR:
            if (!r->linear) return fail(PRH_E_INVALID, "renderer was created without keep_linear");
            src = r->linear, need = 12 * P;
            break;
        case PRH_BUF_SHADOW: {
            need = (size_t)4 * r->n_slots * r->SW * r->SH;
            if (bytes < need) return fail(PRH_E_INVALID, "prh_read_buffer: buffer too small");
            HIP_TRY(hipSetDevice(r->ctx->device));
            if (need) HIP_TRY(hipMemcpyAsync(dst, r->smaps, need, hipMemcpyDeviceToHost, r->ctx->stream));
            HIP_TRY(hipStreamSynchronize(r->ctx->stream));
            uint32_t* u = (uint32_t*)dst;
            for (size_t i = 0; i < need / 4; i++) {  // ordered depth bits -> float32
                float f = depth_from_ord(u[i]);
                memcpy(&u[i], &f, 4);
            }
            return PRH_OK;
        }
        case PRH_BUF_PREGAMMA:
            if (!r->pregamma) return fail(PRH_E_INVALID, "renderer was created without keep_linear");
            src = r->pregamma, need = 4 * P;
            break;
        case PRH_BUF_SHADOW_COUNT: need = P; break;
        default: return fail(PRH_E_INVALID, "unknown buffer id");
    }
    if (bytes < need) return fail(PRH_E_INVALID, "prh_read_buffer: buffer too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    if (which == PRH_BUF_SHADOW_COUNT) {
        if (!r->has_frame) return fail(PRH_E_INVALID, "prh_read_buffer: no rendered frame");
        const uint8_t* c = nullptr;
        int rc = shadow_counts(r, r->ctx->stream, &c);
        if (rc) return rc;
        if (!c) {
            std::memset(dst, 0, need);
            return PRH_OK;
        }
        src = c;
    }
    if (need) HIP_TRY(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, r->ctx->stream));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    return PRH_OK;
}

// The last frame's per-pass times from its events (finish_frame left them).
void fill_times(prh_renderer* r) {
    if (!r->times_pending) return;
    r->times_pending = false;
    prh_frame_stats& S = r->stats;
    float ms[6] = {0};
    for (int i = 0; i < 6; i++) (void)hipEventElapsedTime(&ms[i], r->ev[i], r->ev[i + 1]);
    (void)hipEventElapsedTime(&S.ms_total, r->ev[0], r->ev[6]);
    // (the shadow passes ran beside the main view's: their own span)
    S.ms_shadow = 0.0f;
    if (r->n_slots > 0 && !r->times_nan_rerun) (void)hipEventElapsedTime(&S.ms_shadow, r->ev_s[0], r->ev_s[1]);
    S.ms_vertex = ms[1], S.ms_setup = ms[2], S.ms_raster = ms[3], S.ms_ao = ms[4];
    (void)hipEventElapsedTime(&S.ms_shade, r->ev[7], r->ev[6]);  // k_shade alone
    if (r->split_shade_last) {  // the AO on the second stream; the shading's first half beside it
        (void)hipEventElapsedTime(&S.ms_ao, r->ev_sh[0], r->ev_sh[1]);
        S.ms_shade += ms[4];
    }
}

int prh_last_stats(prh_renderer* r, prh_frame_stats* out) {
    if (!r || !out) return fail(PRH_E_INVALID, "prh_last_stats: null argument");
    HIP_TRY(hipSetDevice(r->ctx->device));
    fill_times(r);
    *out = r->stats;
    return PRH_OK;
}

int prh_diff_sizes(prh_renderer* r, int32_t* n_vertices, int32_t* n_materials, int32_t* n_lights) {
    if (!r || !n_vertices || !n_materials || !n_lights) return fail(PRH_E_INVALID, "prh_diff_sizes: null argument");
    *n_vertices = r->sc->d.n_tverts, *n_materials = r->sc->d.n_materials, *n_lights = r->sc->d.n_lights;
    return PRH_OK;
}

int prh_diff_forward(prh_renderer* r, const prh_diff_params* p, float* linear, size_t bytes) {
    if (!r || !p || !linear) return fail(PRH_E_INVALID, "prh_diff_forward: null argument");
    if (bytes < (size_t)12 * r->W * r->H) return fail(PRH_E_INVALID, "prh_diff_forward: buffer too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    int rc = diff_run(r, p, nullptr, nullptr, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(linear, r->diff_lin, (size_t)12 * r->W * r->H, hipMemcpyDeviceToHost, r->ctx->stream));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    return PRH_OK;
}

int prh_diff_backward(prh_renderer* r, const prh_diff_params* p, const float* target, size_t bytes, double* loss,
                      const prh_diff_grads* grads) {
    if (!r || !p || !target || !loss || !grads) return fail(PRH_E_INVALID, "prh_diff_backward: null argument");
    if (bytes < (size_t)12 * r->W * r->H) return fail(PRH_E_INVALID, "prh_diff_backward: target too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    return diff_run(r, p, target, loss, grads);
}

int prh_diff_backward_noalloc(prh_renderer* r, const prh_diff_params* p, const float* target, size_t bytes,
                              double* loss, const prh_diff_grads* grads) {
    if (!r || !p || !target || !loss || !grads) return fail(PRH_E_INVALID, "prh_diff_backward_noalloc: null argument");
    const size_t P = (size_t)r->W * r->H;
    if (bytes < 12 * P) return fail(PRH_E_INVALID, "prh_diff_backward_noalloc: target too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    if (!r->has_frame) return fail(PRH_E_INVALID, "differentiable shading needs a rendered frame (prh_render)");
    int rc = check_params(r, p);
    if (rc) return rc;
    uint8_t* base = r->ctx->scratch;
    if (!base) return fail(PRH_E_HIP, "prh_diff_backward_noalloc: no scratch array on this device: " + r->ctx->scratch_err);
    // the scratch array's layout for this frame, each part 256-byte aligned
    const DiffLayout l = diff_layout(r);
    const bool counts = r->sh.shadow_map;
    size_t need = 0;
    auto take = [&](size_t n) {
        const size_t o = need;
        need += (n + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_par = take(4 * l.total), o_grad = take(4 * l.total), o_tgt = take(12 * P), o_loss = take(8);
    const size_t o_cnt = counts ? take(P) : 0;
    if (need > PRH_NOALLOC_SCRATCH_BYTES)
        return fail(PRH_E_INVALID, "prh_diff_backward_noalloc: the frame needs " + std::to_string(need) +
                                       " bytes of scratch, more than the " +
                                       std::to_string(PRH_NOALLOC_SCRATCH_BYTES) + " available");
    float* par = reinterpret_cast<float*>(base + o_par);
    float* grad = reinterpret_cast<float*>(base + o_grad);
    float* tgt = reinterpret_cast<float*>(base + o_tgt);
    double* lossd = reinterpret_cast<double*>(base + o_loss);
    uint8_t* cnt = counts ? base + o_cnt : nullptr;
    const DevScene& d = r->sc->d;
    const size_t V = d.n_tverts, M = d.n_materials, L = d.n_lights;
    hipStream_t st = r->ctx->stream;
    if ((rc = upload_params(r, l, p, par, st))) return rc;
    HIP_TRY(hipMemcpyAsync(tgt, target, 12 * P, hipMemcpyHostToDevice, st));
    // the array keeps the previous call's contents: clear what accumulates,
    // with fill kernels (no memset nodes if a caller captures this)
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(grad), l.total, 0u);
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(lossd), 1, 0ull);
    if (counts) launch_shadow_count(st, d, r->sh, shade_buffers(r), cnt);  // every pixel written
    HIP_TRY(hipGetLastError());
    DiffBuffers b = diff_buffers(r, l, par, grad, lossd, p->normal != nullptr, cnt);
    b.target = tgt, b.linear = nullptr;
    launch_diff(st, true, d, r->vc, b);
    HIP_TRY(hipGetLastError());
    auto down = [&](float* dst, size_t off, size_t n) -> hipError_t {
        return dst && n ? hipMemcpyAsync(dst, grad + off, n * 4, hipMemcpyDeviceToHost, st) : hipSuccess;
    };
    HIP_TRY(down(grads->albedo, l.albedo, 3 * V));
    HIP_TRY(down(grads->normal, l.normal, 3 * V));
    HIP_TRY(down(grads->kd, l.kd, 3 * M));
    HIP_TRY(down(grads->ks, l.ks, 3 * M));
    HIP_TRY(down(grads->shininess, l.shin, M));
    HIP_TRY(down(grads->light_pos, l.lpos, 3 * L));
    HIP_TRY(down(grads->light_int, l.lint, L));
    HIP_TRY(hipMemcpyAsync(loss, lossd, 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_fit_new(prh_renderer* const* views, int32_t n, int32_t flags, prh_fit** out) {
    if (!views || n <= 0 || !out) return fail(PRH_E_INVALID, "prh_fit_new: no views");
    *out = nullptr;
    if (flags & ~PRH_FIT_STORED_GBUFFER) return fail(PRH_E_INVALID, "prh_fit_new: unknown flags");
    prh_scene* sc = views[0] ? views[0]->sc : nullptr;
    for (int32_t i = 0; i < n; i++) {
        if (!views[i] || views[i]->sc != sc) return fail(PRH_E_INVALID, "prh_fit_new: views of one scene");
        if (!views[i]->has_frame) return fail(PRH_E_INVALID, "prh_fit_new: every view needs a rendered frame");
    }
    if (sc->d.n_lights > DIFF_MAX_LIGHTS)
        return fail(PRH_E_UNSUPPORTED, "differentiable shading takes at most 8 lights");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    hipStream_t st = sc->ctx->stream;
    prh_fit* f = new prh_fit();
    f->ctx = sc->ctx, f->sc = sc, f->stored = (flags & PRH_FIT_STORED_GBUFFER) != 0;
    f->views.assign(views, views + n);
    std::sort(f->views.begin(), f->views.end());
    if (std::adjacent_find(f->views.begin(), f->views.end()) != f->views.end()) {
        delete f;
        return fail(PRH_E_INVALID, "prh_fit_new: a view appears twice");
    }
    f->views.assign(views, views + n);
    for (prh_renderer* r : f->views) r->live_fits++;
    f->l = diff_layout(views[0]);
    const size_t T = f->l.total;
    int rc = 0;
    auto chk = [&](int v) {
        if (!rc) rc = v;
    };
    chk(falloc(f, &f->par, T));
    chk(falloc(f, &f->grad, T));
    chk(falloc(f, &f->base, T));
    chk(falloc(f, &f->theta, T));
    chk(falloc(f, &f->m, T));
    chk(falloc(f, &f->v, T));
    chk(falloc(f, &f->lower, T));
    chk(falloc(f, &f->idx, T));
    chk(falloc(f, &f->loss, 1));
    chk(falloc(f, &f->step, 1));
    chk(falloc(f, &f->d_views, (size_t)n));
    for (int k = 0; k < 4 && !rc; k++)
        if (hipEventCreate(&f->ev[k]) != hipSuccess) rc = fail(PRH_E_HIP, "hipEventCreate failed");
    double vis_bytes = 0;
    for (int32_t i = 0; i < n && !rc; i++) {
        prh_renderer* r = views[i];
        const size_t P = (size_t)r->W * r->H;
        float* t = nullptr;
        chk(falloc(f, &t, 3 * P));
        f->target.push_back(t);
        const uint8_t* c = nullptr;
        if (!rc) chk(shadow_counts(r, st, &c));
        f->count.push_back(c);
        float4* g = nullptr;
        if (f->stored && !rc) chk(falloc(f, &g, 4 * P));
        f->gbuf.push_back(g);
        f->serials.push_back(r->frame_serial);
        vis_bytes += f->stored ? 64.0 * (double)P : (8.0 + (c ? 1.0 : 0.0)) * (double)P;
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(PRH_E_HIP, "prh_fit_new: shadow counts failed");
    if (rc) {
        prh_fit_free(f);
        return rc;
    }
    f->stats = prh_fit_stats{};
    f->stats.visibility_bytes = vis_bytes;
    f->stats.n_views = n;
    *out = f;
    return PRH_OK;
}

int prh_fit_free(prh_fit* f) {
    if (!f) return PRH_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    for (int k = 0; k < 4; k++)
        if (f->ev[k]) (void)hipEventDestroy(f->ev[k]);
    free_list(f->allocs);
    for (prh_renderer* r : f->views) r->live_fits--;
    delete f;
    return PRH_OK;
}

int prh_fit_targets(prh_fit* f, const prh_diff_params* truth) {
    if (!f || !truth) return fail(PRH_E_INVALID, "prh_fit_targets: null argument");
    if (truth->normal) return fail(PRH_E_UNSUPPORTED, "prh_fit: the views' own normals only");
    int rc = fit_check(f);
    if (!rc) rc = check_params(f->views[0], truth);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t st = f->ctx->stream;
    if ((rc = upload_params(f->views[0], f->l, truth, f->par, st))) return rc;
    if ((rc = fit_views(f, false, DIFF_ALL, st))) return rc;
    launch_diff_views(st, false, f->sc->d, f->d_views, (int32_t)f->views.size(), f->total_blocks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_fit_set_target(prh_fit* f, int32_t view, const float* target, size_t bytes) {
    if (!f || !target || view < 0 || view >= (int32_t)f->views.size())
        return fail(PRH_E_INVALID, "prh_fit_set_target: bad argument");
    const prh_renderer* r = f->views[view];
    const size_t need = (size_t)12 * r->W * r->H;
    if (bytes < need) return fail(PRH_E_INVALID, "prh_fit_set_target: target too small");
    HIP_TRY(hipSetDevice(f->ctx->device));
    HIP_TRY(hipMemcpyAsync(f->target[view], target, need, hipMemcpyHostToDevice, f->ctx->stream));
    HIP_TRY(hipStreamSynchronize(f->ctx->stream));
    return PRH_OK;
}

int prh_fit_run(prh_fit* f, const prh_diff_params* start, const prh_fit_options* o, const prh_diff_grads* result,
                double* loss_history) {
    if (!f || !start || !o) return fail(PRH_E_INVALID, "prh_fit_run: null argument");
    const int32_t known = PRH_FIT_ALBEDO | PRH_FIT_KD | PRH_FIT_KS | PRH_FIT_SHININESS | PRH_FIT_LIGHT_POS |
                          PRH_FIT_LIGHT_INT;
    if (o->steps < 0 || (o->groups & ~known)) return fail(PRH_E_INVALID, "prh_fit_run: bad steps or parameter groups");
    if (!(o->lr > 0.0f) || !(o->beta1 >= 0.0f && o->beta1 < 1.0f) || !(o->beta2 >= 0.0f && o->beta2 < 1.0f) ||
        !(o->eps > 0.0f))
        return fail(PRH_E_INVALID, "prh_fit_run: bad Adam constants");
    if (f->stored && (o->groups & PRH_FIT_ALBEDO))
        return fail(PRH_E_UNSUPPORTED, "prh_fit_run: albedo gradients need the rebuilt G-buffer");
    if (start->normal) return fail(PRH_E_UNSUPPORTED, "prh_fit: the views' own normals only");
    int rc = fit_check(f);
    if (!rc) rc = check_params(f->views[0], start);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t st = f->ctx->stream;
    const DiffLayout& l = f->l;
    // parameters and the Adam state
    if ((rc = upload_params(f->views[0], l, start, f->par, st))) return rc;
    HIP_TRY(hipMemcpyAsync(f->base, f->par, l.total * 4, hipMemcpyDeviceToDevice, st));
    std::vector<int32_t> idx;
    std::vector<float> lower;
    fit_groups(f, o->groups, idx, lower);
    const int32_t n_idx = (int32_t)idx.size();
    if (n_idx) {
        HIP_TRY(hipMemcpyAsync(f->idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(f->lower, lower.data(), lower.size() * 4, hipMemcpyHostToDevice, st));
    }
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(f->theta), l.total, 0u);
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(f->m), l.total, 0u);
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(f->v), l.total, 0u);
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(f->step), 1, 0u);
    HIP_TRY(hipStreamSynchronize(st));  // idx and lower are locals
    if (o->steps > f->hist_cap) {
        if (f->hist) {
            auto it = std::find(f->allocs.begin(), f->allocs.end(), (void*)f->hist);
            if (it != f->allocs.end()) f->allocs.erase(it);
            (void)hipFree(f->hist);
            f->hist = nullptr, f->hist_cap = 0;
        }
        if ((rc = falloc(f, &f->hist, (size_t)o->steps))) return rc;
        f->hist_cap = o->steps;
    }
    // the stored G-buffer, from the start parameters
    HIP_TRY(hipEventRecord(f->ev[2], st));
    if (f->stored) {
        for (size_t i = 0; i < f->views.size(); i++) {
            prh_renderer* r = f->views[i];
            const DiffBuffers b = diff_buffers(r, l, f->par, f->grad, f->loss, false, f->count[i]);
            launch_diff_store_gbuffer(st, f->sc->d, r->vc, b, f->gbuf[i]);
        }
    }
    HIP_TRY(hipEventRecord(f->ev[3], st));
    if ((rc = fit_views(f, true, (uint32_t)o->groups, st))) return rc;
    HIP_TRY(hipGetLastError());
    const AdamConsts a{o->lr, o->beta1, o->beta2, o->eps, n_idx, f->hist_cap};
    // one step: clear, adjoint over every view, Adam; captured once, replayed
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (o->steps > 0) {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        launch_fill_u32(st, reinterpret_cast<uint32_t*>(f->grad), l.total, 0u);
        launch_fill_u64(st, reinterpret_cast<uint64_t*>(f->loss), 1, 0ull);
        launch_diff_views(st, true, f->sc->d, f->d_views, (int32_t)f->views.size(), f->total_blocks);
        launch_adam(st, f->par, f->grad, f->base, f->theta, f->m, f->v, f->idx, f->lower, a, f->step, f->loss,
                    f->hist);
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (e != hipSuccess) return fail(PRH_E_HIP, std::string("prh_fit_run capture: ") + hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(PRH_E_HIP, std::string("prh_fit_run instantiate: ") + hipGetErrorString(e2));
        }
    }
    hipError_t e = hipEventRecord(f->ev[0], st);
    for (int32_t k = 0; k < o->steps && e == hipSuccess; k++) e = hipGraphLaunch(ge, st);
    if (e == hipSuccess) e = hipEventRecord(f->ev[1], st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(PRH_E_HIP, std::string("prh_fit_run: ") + hipGetErrorString(e));
    f->stats.steps = o->steps;
    (void)hipEventElapsedTime(&f->stats.ms_steps, f->ev[0], f->ev[1]);
    f->stats.ms_gbuffer = 0.0f;
    if (f->stored) (void)hipEventElapsedTime(&f->stats.ms_gbuffer, f->ev[2], f->ev[3]);
    if (loss_history && o->steps > 0)
        HIP_TRY(hipMemcpy(loss_history, f->hist, (size_t)o->steps * 8, hipMemcpyDeviceToHost));
    if (result) {
        const DevScene& d = f->sc->d;
        const size_t V = d.n_tverts, M = d.n_materials, L = d.n_lights;
        auto down = [&](float* dst, size_t off, size_t n) -> hipError_t {
            return dst && n ? hipMemcpy(dst, f->par + off, n * 4, hipMemcpyDeviceToHost) : hipSuccess;
        };
        HIP_TRY(down(result->albedo, l.albedo, 3 * V));
        HIP_TRY(down(result->kd, l.kd, 3 * M));
        HIP_TRY(down(result->ks, l.ks, 3 * M));
        HIP_TRY(down(result->shininess, l.shin, M));
        HIP_TRY(down(result->light_pos, l.lpos, 3 * L));
        HIP_TRY(down(result->light_int, l.lint, L));
    }
    return PRH_OK;
}

int prh_fit_last_stats(prh_fit* f, prh_fit_stats* out) {
    if (!f || !out) return fail(PRH_E_INVALID, "prh_fit_last_stats: null argument");
    *out = f->stats;
    return PRH_OK;
}


// ---- headless present (gpu/surface.go:1-183)
int prh_surface_new(prh_ctx* ctx, int32_t width, int32_t height, int32_t frames, prh_surface** out) {
    if (!ctx || !out) return fail(PRH_E_INVALID, "prh_surface_new: null argument");
    *out = nullptr;
    if (width <= 0 || height <= 0) return fail(PRH_E_INVALID, "prh_surface_new: surface size must be > 0");
    if (frames <= 0) frames = 2;  // SurfaceDescriptor.Frames: double buffering by default
    if (frames > 16) return fail(PRH_E_INVALID, "prh_surface_new: at most 16 frames");
    HIP_TRY(hipSetDevice(ctx->device));
    prh_surface* s = new prh_surface();
    s->ctx = ctx, s->frames = frames;
    const int rc = surface_alloc(s, width, height);
    if (rc) {
        delete s;
        return rc;
    }
    ctx->live_surfaces++;
    *out = s;
    return PRH_OK;
}

int prh_surface_free(prh_surface* s) {
    if (!s) return PRH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    surface_release(s);
    s->ctx->live_surfaces--;
    delete s;
    return PRH_OK;
}

int prh_surface_acquire(prh_surface* s, int32_t* index) {
    if (!s) return fail(PRH_E_INVALID, "prh_surface_acquire: null surface");
    s->acquired = true;
    if (index) *index = s->frame % s->frames;
    return PRH_OK;
}

int prh_surface_render(prh_surface* s, prh_renderer* r) {
    if (!s || !r) return fail(PRH_E_INVALID, "prh_surface_render: null argument");
    if (!s->acquired) return fail(PRH_E_INVALID, "prh_surface_render: no acquired texture");
    if (r->ctx != s->ctx) return fail(PRH_E_INVALID, "prh_surface_render: renderer of another context");
    if (r->OW != s->w || r->OH != s->h) return fail(PRH_E_INVALID, "prh_surface_render: renderer and surface sizes differ");
    r->present_to = s->tex[s->frame % s->frames];
    const int rc = prh_render(r);
    r->present_to = nullptr;
    return rc;
}

int prh_surface_write(prh_surface* s, const uint8_t* rgba, size_t bytes) {
    if (!s || !rgba) return fail(PRH_E_INVALID, "prh_surface_write: null argument");
    if (!s->acquired) return fail(PRH_E_INVALID, "prh_surface_write: no acquired texture");
    const size_t need = (size_t)4 * s->w * s->h;
    if (bytes < need) return fail(PRH_E_INVALID, "prh_surface_write: image too small");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipMemcpyAsync(s->tex[s->frame % s->frames], rgba, need, hipMemcpyHostToDevice, s->ctx->stream));
    return PRH_OK;
}

int prh_surface_present(prh_surface* s) {
    if (!s) return fail(PRH_E_INVALID, "prh_surface_present: null surface");
    if (!s->acquired) return fail(PRH_E_INVALID, "prh_surface_present: Present without AcquireNextTexture");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));  // headless: the frame is complete
    s->acquired = false;
    s->frame++;
    return PRH_OK;
}

int prh_surface_read(prh_surface* s, uint8_t* rgba, size_t bytes) {
    if (!s || !rgba) return fail(PRH_E_INVALID, "prh_surface_read: null argument");
    if (s->frame == 0) return fail(PRH_E_INVALID, "prh_surface_read: nothing presented yet");
    const size_t need = (size_t)4 * s->w * s->h;
    if (bytes < need) return fail(PRH_E_INVALID, "prh_surface_read: buffer too small");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipMemcpyAsync(rgba, s->tex[(s->frame - 1) % s->frames], need, hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    return PRH_OK;
}

int prh_surface_resize(prh_surface* s, int32_t width, int32_t height) {
    if (!s) return fail(PRH_E_INVALID, "prh_surface_resize: null surface");
    if (width <= 0 || height <= 0) return fail(PRH_E_INVALID, "prh_surface_resize: surface size must be > 0");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    surface_release(s);
    s->frame = 0, s->acquired = false;
    return surface_alloc(s, width, height);
}

int prh_surface_size(prh_surface* s, int32_t* width, int32_t* height) {
    if (!s || !width || !height) return fail(PRH_E_INVALID, "prh_surface_size: null argument");
    *width = s->w, *height = s->h;
    return PRH_OK;
}


// ---- autotuning
int prh_get_tuning(prh_renderer* r, prh_tuning* out) {
    if (!r || !out) return fail(PRH_E_INVALID, "prh_get_tuning: null argument");
    *out = r->tune;
    return PRH_OK;
}

int prh_set_tuning(prh_renderer* r, const prh_tuning* t) {
    if (!r || !t) return fail(PRH_E_INVALID, "prh_set_tuning: null argument");
    const bool pow2 = t->item_prims > 0 && (t->item_prims & (t->item_prims - 1)) == 0;
    if (t->raster_wg_per_cu < 1 || t->raster_wg_per_cu > 16 || !pow2 || t->item_prims < 64 ||
        t->item_prims > 8192 || t->shadow_direct_max < 0 || t->shadow_direct_max > 256 || t->main_direct_max < 0 ||
        t->main_direct_max > 256 || (t->overlap_shadows != 0 && t->overlap_shadows != 1) ||
        (t->raster_tile != 8 && t->raster_tile != 16 && t->raster_tile != 32))
        return fail(PRH_E_INVALID, "prh_set_tuning: a knob out of range");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    const bool realloc = t->item_prims != r->tune.item_prims;
    r->tune.raster_wg_per_cu = t->raster_wg_per_cu, r->tune.item_prims = t->item_prims;
    r->tune.shadow_direct_max = t->shadow_direct_max, r->tune.main_direct_max = t->main_direct_max;
    r->tune.overlap_shadows = t->overlap_shadows;
    r->tune.raster_tile = t->raster_tile;
    r->shc.direct_max = t->shadow_direct_max;
    raster_grid(r);
    if (realloc) {  // the work-item table is sized by item_prims
        free_work(r);
        int rc = alloc_work(r);
        if (rc) return rc;
    }
    r->work_gen++;  // a cached batch recorded the old values
    return PRH_OK;
}

int prh_autotune(prh_renderer* r, int32_t frames, prh_tuning* out) {
    if (!r) return fail(PRH_E_INVALID, "prh_autotune: null renderer");
    if (frames < 1) frames = 5;
    auto time = [&](const prh_tuning& t, float* ms) -> int {
        int rc = prh_set_tuning(r, &t);
        if (rc) return rc;
        if ((rc = prh_render(r))) return rc;  // warm-up (and any regrowth)
        std::vector<float> v;
        for (int k = 0; k < frames; k++) {
            if ((rc = prh_render(r))) return rc;
            fill_times(r);
            v.push_back(r->stats.ms_total);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        return PRH_OK;
    };
    static const int32_t c_ip[] = {256, 512, 1024, 2048, 4096};
    static const int32_t c_wg[] = {1, 2, 4, 8};
    static const int32_t c_sd[] = {0, 16, 32, 64, 128, 256};
    static const int32_t c_md[] = {0, 24, 48, 96, 192};
    static const int32_t c_ov[] = {0, 1};
    static const int32_t c_rt[] = {8, 16, 32};
    const prh_tuning start = r->tune;
    prh_tuning best = start;
    float best_ms = 0.0f;
    int rc = time(best, &best_ms);
    if (rc) return rc;
    // coordinate descent: each knob over its candidates, the others at their best
    for (int knob = 0; knob < 6; knob++) {
        const int32_t* c = knob == 0 ? c_ip : knob == 1 ? c_wg : knob == 2 ? c_sd : knob == 3 ? c_md : knob == 4 ? c_ov : c_rt;
        const int n = knob == 0 ? 5 : knob == 1 ? 4 : knob == 2 ? 6 : knob == 3 ? 5 : knob == 4 ? 2 : 3;
        for (int k = 0; k < n; k++) {
            prh_tuning t = best;
            int32_t* f = knob == 0 ? &t.item_prims
                       : knob == 1 ? &t.raster_wg_per_cu
                       : knob == 2 ? &t.shadow_direct_max
                       : knob == 3 ? &t.main_direct_max
                       : knob == 4 ? &t.overlap_shadows : &t.raster_tile;
            if (*f == c[k]) continue;
            *f = c[k];
            float ms = 0.0f;
            if ((rc = time(t, &ms))) return rc;
            if (ms < best_ms) best = t, best_ms = ms;
        }
    }
    // a second look at the winner against the starting point, over twice the
    // frames, so that one noisy median does not move the knobs
    if (std::memcmp(&best, &start, offsetof(prh_tuning, ms)) != 0) {
        frames *= 2;
        float ms_best = 0.0f, ms_start = 0.0f;
        if ((rc = time(start, &ms_start)) || (rc = time(best, &ms_best))) return rc;
        if (ms_start <= ms_best) best = start, best_ms = ms_start;
        else best_ms = ms_best;
    }
    best.ms = best_ms;
    if ((rc = prh_set_tuning(r, &best))) return rc;
    r->tune.ms = best_ms;
    if ((rc = prh_render(r))) return rc;  // leave a frame of the chosen configuration
    if (out) *out = r->tune;
    return PRH_OK;
}

int prh_autotune_batch(prh_renderer* const* rs, int32_t n, int32_t frames, prh_tuning out[2]) {
    if (!rs || n <= 0) return fail(PRH_E_INVALID, "prh_autotune_batch: no renderers");
    for (int i = 0; i < n; i++)
        if (!rs[i]) return fail(PRH_E_INVALID, "prh_autotune_batch: null renderer");
    if (frames < 1) frames = 5;
    std::vector<int> cls(n);
    for (int i = 0; i < n; i++) cls[i] = (int64_t)rs[i]->W * rs[i]->H >= (1 << 19) ? 1 : 0;
    std::vector<prh_tuning> start(n);
    for (int i = 0; i < n; i++) start[i] = rs[i]->tune;
    // the classes start from the knobs of their first view
    prh_tuning cur[2];
    for (int c = 0; c < 2; c++) {
        cur[c] = rs[0]->tune;
        for (int i = 0; i < n; i++)
            if (cls[i] == c) {
                cur[c] = rs[i]->tune;
                break;
            }
    }
    auto apply = [&](const prh_tuning* t) -> int {
        for (int i = 0; i < n; i++) {
            prh_tuning v = t[cls[i]];
            v.overlap_shadows = rs[i]->tune.overlap_shadows;  // (not used by a wide batch)
            int rc = prh_set_tuning(rs[i], &v);
            if (rc) return rc;
        }
        return PRH_OK;
    };
    auto time = [&](const prh_tuning* t, int reps, float* ms) -> int {
        int rc = apply(t);
        if (rc) return rc;
        float m = 0.0f;
        if ((rc = prh_render_batch(rs, n, 0, &m))) return rc;  // warm-up, regrowth, tables
        std::vector<float> v;
        for (int k = 0; k < reps; k++) {
            if ((rc = prh_render_batch(rs, n, 0, &m))) return rc;
            v.push_back(m);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        return PRH_OK;
    };
    static const int32_t c_ip[] = {256, 512, 1024, 2048, 4096};
    static const int32_t c_wg[] = {2, 4, 8, 16};
    static const int32_t c_sd[] = {0, 16, 32, 64, 128, 256};
    static const int32_t c_md[] = {0, 24, 48, 96, 192};
    static const int32_t c_rt[] = {8, 16, 32};
    const bool has[2] = {std::count(cls.begin(), cls.end(), 0) > 0, std::count(cls.begin(), cls.end(), 1) > 0};
    const prh_tuning first[2] = {cur[0], cur[1]};
    float best_ms = 0.0f;
    int rc = time(cur, frames, &best_ms);
    if (rc) return rc;
    for (int knob = 0; knob < 5; knob++)
        for (int c = 0; c < 2; c++) {
            if (!has[c]) continue;
            const int32_t* cand = knob == 0 ? c_ip : knob == 1 ? c_wg : knob == 2 ? c_sd : knob == 3 ? c_md : c_rt;
            const int nc = knob == 0 ? 5 : knob == 1 ? 4 : knob == 2 ? 6 : knob == 3 ? 5 : 3;
            for (int k = 0; k < nc; k++) {
                prh_tuning t[2] = {cur[0], cur[1]};
                int32_t* f = knob == 0 ? &t[c].item_prims
                           : knob == 1 ? &t[c].raster_wg_per_cu
                           : knob == 2 ? &t[c].shadow_direct_max
                           : knob == 3 ? &t[c].main_direct_max : &t[c].raster_tile;
                if (*f == cand[k]) continue;
                *f = cand[k];
                float ms = 0.0f;
                if ((rc = time(t, frames, &ms))) return rc;
                if (ms < best_ms) cur[0] = t[0], cur[1] = t[1], best_ms = ms;
            }
        }
    if (std::memcmp(&cur[0], &first[0], offsetof(prh_tuning, ms)) != 0 ||
        std::memcmp(&cur[1], &first[1], offsetof(prh_tuning, ms)) != 0) {
        float ms_first = 0.0f, ms_best = 0.0f;
        if ((rc = time(first, 2 * frames, &ms_first)) || (rc = time(cur, 2 * frames, &ms_best))) return rc;
        if (ms_first <= ms_best) cur[0] = first[0], cur[1] = first[1], best_ms = ms_first;
        else best_ms = ms_best;
    }
    cur[0].ms = cur[1].ms = best_ms;
    if ((rc = apply(cur))) return rc;
    for (int i = 0; i < n; i++) rs[i]->tune.ms = best_ms;
    float m = 0.0f;
    if ((rc = prh_render_batch(rs, n, 0, &m))) return rc;  // leave a frame of the chosen knobs in every view
    if (out) out[0] = cur[0], out[1] = cur[1];
    return PRH_OK;
}

int prh_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0) return fail(PRH_E_INVALID, "prh_host_alloc: bad argument");
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(PRH_E_NOMEM, "prh_host_alloc: hipHostMalloc failed");
    }
    return PRH_OK;
}

int prh_host_free(void* p) {
    if (p) HIP_TRY(hipHostFree(p));
    return PRH_OK;
}

}  // extern "C"
