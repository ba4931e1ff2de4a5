// diff.hip — the differentiable shading pass of libprh: forward and adjoint
// of the deferred Blinn-Phong shading at fixed visibility.
//
// polyred has no derivative path; this extends its shading model
// (shader/blinn_cpu.go:24-106, the sums before Round) with real-valued
// parameters:
//   albedo     per transformed vertex, interpolated perspective-correctly as
//              drawClipped interpolates the vertex colour (render/raster.go:
//              500-507); it stands where FragmentShader reads the uniform
//              texture colour (ambient and diffuse terms);
//   normal     per transformed vertex, unit(sum bc_k N_k) as drawClipped's
//              normal (render/raster.go:509-511);
//   Kd, Ks     per material (RGB, the 0..255 scale of BlinnPhong.Diffuse and
//              .Specular), shininess per material;
//   position and intensity of every light.
// A pixel's geometry is rebuilt from its visibility key as k_shade does (a
// clipped piece takes its parent triangle); the world position, view vector
// and light colours are fixed.  The loss is sum over shaded pixels and RGB of
// (lin - target)^2.  Gradients are accumulated with one atomic per distinct
// triangle / material and per light in each wave.
//
// Shadows enter at fixed visibility, as the reference applies them after the
// shading (render/raster.go:339-347): the pixel's sums are scaled by
// 2^-count, count the lights for which shadingVisibility reported true in the
// frame that fixed the visibility.  A view's pixel state is either rebuilt
// from the visibility buffer at every launch or read from a stored G-buffer.
#include "kernels.hpp"

using namespace gm;

namespace prh {

namespace {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(F3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ F3 ld3(const float* p) { return f3(p[0], p[1], p[2]); }
__device__ __forceinline__ F3 xyz(float4 a) { return f3(a.x, a.y, a.z); }
// gradient of u = v / |v| pulled back to v: (g - u (u.g)) / |v|
__device__ __forceinline__ F3 unit_back(F3 u, float l, F3 g) { return (1.0f / l) * (g - dot(u, g) * u); }

__device__ __forceinline__ float wsum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ double wsum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// What one pixel needs for the shading and its adjoint.
struct DPix {
    bool act;
    int g, mat, vid[3];
    float w[3], bc[3];  // perspective weights (albedo) and barycentrics (normal)
    F3 a, nraw, n, x, V;
    float nl;
    float sw;  // shadow scale 2^-count
    bool flat;
};

// G-buffer planes: (x, material), (n, shadow scale), (V, 0), (albedo, 0).
__device__ DPix stored_pixel(const ViewConsts& v, const DiffBuffers& b, int x, int y, bool inb) {
    DPix p;
    p.act = false;
    if (!inb) return p;
    const size_t P = (size_t)v.W * v.H, o = (size_t)y * v.W + x;
    const float4 g0 = b.gbuf[o];
    const int mat = __float_as_int(g0.w);
    if (mat < 0) return p;
    const float4 g1 = b.gbuf[P + o], g2 = b.gbuf[2 * P + o], g3 = b.gbuf[3 * P + o];
    p.act = true;
    p.g = -1, p.mat = mat;
    p.x = xyz(g0), p.n = xyz(g1), p.sw = g1.w, p.V = xyz(g2), p.a = xyz(g3);
    p.nraw = p.n, p.nl = 1.0f, p.flat = true;  // no normal gradient from a stored pixel
    for (int k = 0; k < 3; k++) p.vid[k] = 0, p.w[k] = p.bc[k] = 0.0f;
    return p;
}

__device__ DPix diff_pixel(const DevScene& s, const ViewConsts& v, const DiffBuffers& b, int x, int y, bool inb) {
    if (b.gbuf) return stored_pixel(v, b, x, y, inb);
    DPix p;
    p.act = false;
    if (!inb) return p;
    const uint64_t key = b.vis[(size_t)y * v.W + x];
    if (key == VIS_EMPTY) return p;
    const int g = (int)((uint32_t)key >> 3);
    const int o = s.tri_obj[g];
    const int mesh = s.obj_mesh[o];
    const int t = s.mesh_tri_off[mesh] + (g - s.obj_tri_start[o]);
    int mat = s.tri_mat[t];
    if (mat >= 0) mat += s.obj_mat_base[o];
    if (mat < 0 || mat >= s.n_materials) return p;  // vertex colour: nothing to differentiate
    p.act = true;
    p.g = g;
    p.mat = mat;
    const int vb = s.obj_vtx_start[o];
    float4 P[3];
    for (int k = 0; k < 3; k++) {
        p.vid[k] = vb + s.tri_v[3 * t + k];
        P[k] = b.tv_scr[p.vid[k]];
    }
    const TriSetup ts = tri_setup(P[0].x, P[0].y, P[1].x, P[1].y, P[2].x, P[2].y);
    barycoord(ts, (float)x + 0.5f, (float)y + 0.5f, p.bc);
    const float wc[3] = {P[0].w * p.bc[0], P[1].w * p.bc[1], P[2].w * p.bc[2]};
    const float norm = v.perspect ? 1 / (wc[0] + wc[1] + wc[2]) : 1.0f;
    p.a = f3(0, 0, 0), p.nraw = f3(0, 0, 0), p.x = f3(0, 0, 0);
    F3 m[3];
    for (int k = 0; k < 3; k++) {
        p.w[k] = wc[k] * norm;
        p.a = p.a + p.w[k] * ld3(b.albedo + 3 * (size_t)p.vid[k]);
        const F3 N = b.normal ? ld3(b.normal + 3 * (size_t)p.vid[k]) : xyz(b.tv_nor[p.vid[k]]);
        p.nraw = p.nraw + p.bc[k] * N;
        m[k] = xyz(b.tv_wp[p.vid[k]]);
        p.x = p.x + p.bc[k] * m[k];
    }
    p.flat = s.mats[mat].flat_shading != 0;
    if (p.flat) {  // the face normal: fixed
        const F3 e1 = m[1] - m[0], e2 = m[2] - m[0];
        const F3 c = f3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
        p.nl = 1.0f;
        p.n = (1.0f / len3(c)) * c;
    } else {
        p.nl = len3(p.nraw);
        p.n = (1.0f / p.nl) * p.nraw;
    }
    const F3 vv = f3(v.cam[0], v.cam[1], v.cam[2]) - p.x;
    p.V = (1.0f / len3(vv)) * vv;
    p.sw = b.shadow_count ? __builtin_ldexpf(1.0f, -(int)b.shadow_count[(size_t)y * v.W + x]) : 1.0f;
    return p;
}

// One light's terms at a pixel.
struct DLight {
    F3 L, H, d;
    float I, dl, hl, nL, nH, Ld, Ls;
};
__device__ __forceinline__ DLight light_terms(const DevLight& l, const DiffBuffers& b, int li, const DPix& p,
                                              float shin) {
    DLight r;
    if (l.kind == 0) {
        r.d = ld3(b.light_pos + 3 * li) - p.x;
        r.dl = len3(r.d);
        r.L = (1.0f / r.dl) * r.d;
        r.I = b.light_int[li] / r.dl;
    } else {
        r.d = f3(0, 0, 0);
        r.dl = 1.0f;
        r.L = f3(-l.dir[0], -l.dir[1], -l.dir[2]);
        r.I = b.light_int[li];
    }
    const F3 hv = r.L + p.V;
    r.hl = len3(hv);
    r.H = (1.0f / r.hl) * hv;
    r.nL = dot(p.n, r.L);
    r.nH = dot(p.n, r.H);
    r.Ld = fminf(fmaxf(r.nL, 0.0f), 1.0f);
    r.Ls = __builtin_powf(fminf(fmaxf(r.nH, 0.0f), 1.0f), shin);
    return r;
}

}  // namespace

// Per-workgroup gradient accumulators in LDS: the lights' (4 per light)
// and, for up to DIFF_LDS_MATS materials, the materials' (7 each); the
// waves add their reductions here and the workgroup flushes them to the
// global gradients once, so the global atomics to these few addresses are
// per workgroup, not per wave.
constexpr int DIFF_LDS_MATS = 256;
constexpr int DIFF_ACC = 4 * DIFF_MAX_LIGHTS + 7 * DIFF_LDS_MATS;
struct DiffAcc {
    float g[DIFF_ACC];
    double loss;
};

__device__ __forceinline__ void acc_clear(DiffAcc& a) {
    for (int i = threadIdx.x; i < DIFF_ACC; i += blockDim.x) a.g[i] = 0.0f;
    if (threadIdx.x == 0) a.loss = 0.0;
    __syncthreads();
}

// After the workgroup's last tile: every nonzero accumulator to its gradient.
__device__ __forceinline__ void acc_flush(const DevScene& s, const DiffBuffers& b, DiffAcc& a) {
    __syncthreads();
    const int NL = s.n_lights, NM = s.n_materials < DIFF_LDS_MATS ? s.n_materials : DIFF_LDS_MATS;
    for (int i = threadIdx.x; i < 4 * NL; i += blockDim.x) {
        const float v = a.g[i];
        if (v != 0.0f) atomicAdd((i & 3) < 3 ? &b.g_lpos[3 * (i >> 2) + (i & 3)] : &b.g_lint[i >> 2], v);
    }
    for (int i = threadIdx.x; i < 7 * NM; i += blockDim.x) {
        const float v = a.g[4 * DIFF_MAX_LIGHTS + i];
        const int m = i / 7, j = i % 7;
        if (v != 0.0f) atomicAdd(j < 3 ? &b.g_kd[3 * m + j] : (j < 6 ? &b.g_ks[3 * m + j - 3] : &b.g_shin[m]), v);
    }
    if (threadIdx.x == 0 && a.loss != 0.0) atomicAdd(b.loss, a.loss);
}

// One 16x16 tile (`bid` of the view's grid) of the forward or the adjoint;
// the adjoint adds its light and material gradients and its loss to `acc`.
// Block i takes tile i: the dispatcher deals consecutive workgroups
// round-robin to the 8 XCDs, so each XCD gets a share of every part of the
// view (the covered pixels, which are the work, crowd where the ground is).
template <bool BACKWARD>
__device__ __forceinline__ void diff_body(const DevScene& s, const ViewConsts& v, const DiffBuffers& b, int bid,
                                          DiffAcc& acc) {
    const int tiles_x = v.tiles_x, n_tiles = tiles_x * v.tiles_y;
    const int tile = bid;
    if (tile >= n_tiles) return;  // whole workgroup
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int x = (tile % tiles_x) * TILE + (wv & 1) * 8 + (lane & 7);
    const int y = (tile / tiles_x) * TILE + (wv >> 1) * 8 + (lane >> 3);
    const bool inb = x < v.W && y < v.H;
    const DPix p = diff_pixel(s, v, b, x, y, inb);
    const int NL = s.n_lights;
    float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, shin = 1.0f;
    float sLd = 0.0f, sLs[3] = {0, 0, 0}, lin[3] = {0, 0, 0};
    if (p.act) {
        for (int c = 0; c < 3; c++) kd[c] = b.kd[3 * p.mat + c], ks[c] = b.ks[3 * p.mat + c];
        shin = b.shininess[p.mat];
        for (int li = 0; li < NL; li++) {
            const DevLight& l = s.lights[li];
            const DLight t = light_terms(l, b, li, p, shin);
            sLd += t.Ld * t.I;
            for (int c = 0; c < 3; c++) sLs[c] += t.Ls * l.color[c] * t.I;
        }
        const float a[3] = {p.a.x, p.a.y, p.a.z};
        for (int c = 0; c < 3; c++)
            lin[c] = p.sw * (b.ambient * a[c] + kd[c] * sLd * a[c] / 255.0f + ks[c] * sLs[c] / 255.0f);
    }
    if (!BACKWARD) {
        if (inb && b.linear) {
            const size_t o = (size_t)y * v.W + x;
            for (int c = 0; c < 3; c++) b.linear[3 * o + c] = p.act ? lin[c] : __builtin_nanf("");
        }
        return;
    }
    // ---- adjoint (every lane of the wave takes part in the reductions)
    const float a[3] = {p.a.x, p.a.y, p.a.z};
    float gl[3] = {0, 0, 0};
    double l2 = 0.0;
    if (p.act) {
        const size_t o = (size_t)y * v.W + x;
        for (int c = 0; c < 3; c++) {
            const float d = lin[c] - b.target[3 * o + c];
            gl[c] = 2.0f * d * p.sw;  // through the fixed shadow scale
            l2 += (double)d * (double)d;
        }
    }
    float gm_[7];  // dKd (3), dKs (3), dshininess
    float ga[3];
    for (int c = 0; c < 3; c++) {
        gm_[c] = gl[c] * sLd * a[c] / 255.0f;
        gm_[3 + c] = gl[c] * sLs[c] / 255.0f;
        ga[c] = gl[c] * (b.ambient + kd[c] * sLd / 255.0f);
    }
    gm_[6] = 0.0f;
    F3 gn = f3(0, 0, 0);
    float glight[4 * DIFF_MAX_LIGHTS];  // per light: d/dposition (3), d/dintensity
    for (int li = 0; li < DIFF_MAX_LIGHTS; li++) {
        glight[4 * li] = glight[4 * li + 1] = glight[4 * li + 2] = glight[4 * li + 3] = 0.0f;
        if (li >= NL || !p.act) continue;
        const DevLight& l = s.lights[li];
        const DLight t = light_terms(l, b, li, p, shin);
        float gLd = 0, gLs = 0, gI = 0;
        for (int c = 0; c < 3; c++) {
            gLd += gl[c] * kd[c] * a[c] * t.I / 255.0f;
            gLs += gl[c] * ks[c] * l.color[c] * t.I / 255.0f;
            gI += gl[c] * (kd[c] * t.Ld * a[c] + ks[c] * t.Ls * l.color[c]) / 255.0f;
        }
        const float dLd = (t.nL > 0.0f && t.nL < 1.0f) ? 1.0f : 0.0f;
        float dLs = 0.0f;
        if (t.nH > 0.0f && t.nH < 1.0f) {
            dLs = shin * t.Ls / t.nH;  // d(nH^s)/dnH
            gm_[6] += gLs * t.Ls * __builtin_logf(t.nH);
        }
        gn = gn + (gLd * dLd) * t.L + (gLs * dLs) * t.H;
        F3 gL = (gLd * dLd) * p.n;
        gL = gL + unit_back(t.H, t.hl, (gLs * dLs) * p.n);
        if (l.kind == 0) {
            const float i0 = b.light_int[li];
            F3 gd = unit_back(t.L, t.dl, gL) + (-gI * i0 / (t.dl * t.dl * t.dl)) * t.d;
            glight[4 * li] = gd.x, glight[4 * li + 1] = gd.y, glight[4 * li + 2] = gd.z;
            glight[4 * li + 3] = gI / t.dl;
        } else {
            glight[4 * li + 3] = gI;
        }
    }
    const F3 gnr = (p.act && !p.flat) ? unit_back(p.n, p.nl, gn) : f3(0, 0, 0);

    // per light: one reduction over the wave, lane j adds value j
    if (b.want & (DIFF_LIGHT_POS | DIFF_LIGHT_INT)) {
        float mine = 0.0f;
        for (int j = 0; j < 4 * DIFF_MAX_LIGHTS; j++) {
            if (j >= 4 * NL) break;
            const float sj = wsum(glight[j]);
            if (lane == j) mine = sj;
        }
        if (lane < 4 * NL && lane < 4 * DIFF_MAX_LIGHTS && mine != 0.0f) atomicAdd(&acc.g[lane], mine);
    }
    // per material: one reduction per distinct material of the wave
    if (b.want & (DIFF_KD | DIFF_KS | DIFF_SHININESS)) {
        uint64_t pending = __ballot(p.act);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int lm = __shfl(p.mat, leader);
            const bool same = p.act && p.mat == lm;
            float mine = 0.0f;
            for (int j = 0; j < 7; j++) {
                const float sj = wsum(same ? gm_[j] : 0.0f);
                if (lane == j) mine = sj;
            }
            if (lm < DIFF_LDS_MATS) {
                if (lane < 7 && mine != 0.0f) atomicAdd(&acc.g[4 * DIFF_MAX_LIGHTS + 7 * lm + lane], mine);
            } else if (lane < 3) {
                atomicAdd(&b.g_kd[3 * lm + lane], mine);
            } else if (lane < 6) {
                atomicAdd(&b.g_ks[3 * lm + lane - 3], mine);
            } else if (lane == 6) {
                atomicAdd(&b.g_shin[lm], mine);
            }
            pending &= ~__ballot(same);
        }
    }
    // per triangle: its three vertices' albedo and normal gradients (a
    // stored G-buffer keeps no vertices: the host refuses that combination)
    if (!b.gbuf && (b.want & (DIFF_ALBEDO | DIFF_NORMAL))) {
        const int tri = p.act ? p.g : -1;
        uint64_t pending = __ballot(p.act);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int lt = __shfl(tri, leader);
            const int lv[3] = {__shfl(p.vid[0], leader), __shfl(p.vid[1], leader), __shfl(p.vid[2], leader)};
            const bool same = p.act && tri == lt;
            float mine = 0.0f;
            for (int k = 0; k < 3; k++) {
                const float vals[6] = {p.w[k] * ga[0], p.w[k] * ga[1], p.w[k] * ga[2],
                                       p.bc[k] * gnr.x, p.bc[k] * gnr.y, p.bc[k] * gnr.z};
                for (int j = 0; j < 6; j++) {
                    const float sj = wsum(same ? vals[j] : 0.0f);
                    if (lane == 6 * k + j) mine = sj;
                }
            }
            if (lane < 18) {
                const int k = lane / 6, j = lane % 6;
                const int vk = k == 0 ? lv[0] : (k == 1 ? lv[1] : lv[2]);
                if (j < 3) atomicAdd(&b.g_albedo[3 * (size_t)vk + j], mine);
                else atomicAdd(&b.g_normal[3 * (size_t)vk + j - 3], mine);
            }
            pending &= ~__ballot(same);
        }
    }
    const double tot = wsum(l2);
    if (lane == 0 && tot != 0.0) atomicAdd(&acc.loss, tot);
}

template <bool BACKWARD>
__global__ __launch_bounds__(256) void k_diff(DevScene s, ViewConsts v, DiffBuffers b) {
    __shared__ DiffAcc acc;
    if (BACKWARD) acc_clear(acc);
    diff_body<BACKWARD>(s, v, b, blockIdx.x, acc);
    if (BACKWARD) acc_flush(s, b, acc);
}

// Many views of one scene in one launch: view i owns the tiles
// [block0, block0 + nblocks) of a virtual grid, block0 ascending and a
// multiple of 8 (so a view's tiles keep the XCD spread of its own launch);
// the workgroups stride over that grid, accumulating in LDS, and flush once.
// The views' gradient and loss buffers are the same.
template <bool BACKWARD>
__global__ __launch_bounds__(256) void k_diff_views(DevScene s, const DiffView* __restrict__ views, int32_t n_views,
                                                    int32_t total_blocks) {
    __shared__ DiffAcc acc;
    if (BACKWARD) acc_clear(acc);
    for (int bid = blockIdx.x; bid < total_blocks; bid += gridDim.x) {
        int lo = 0, hi = n_views - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (views[mid].block0 <= bid) lo = mid;
            else hi = mid - 1;
        }
        const DiffView& w = views[lo];
        if (bid - w.block0 < w.nblocks) diff_body<BACKWARD>(s, w.v, w.b, bid - w.block0, acc);
    }
    if (BACKWARD) acc_flush(s, views[0].b, acc);
}

__global__ __launch_bounds__(256) void k_diff_gbuffer(DevScene s, ViewConsts v, DiffBuffers b, float4* __restrict__ g) {
    const size_t P = (size_t)v.W * v.H;
    const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (o >= P) return;
    const int x = (int)(o % (size_t)v.W), y = (int)(o / (size_t)v.W);
    const DPix p = diff_pixel(s, v, b, x, y, true);
    if (!p.act) {
        g[o] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
        return;
    }
    g[o] = make_float4(p.x.x, p.x.y, p.x.z, __int_as_float(p.mat));
    g[P + o] = make_float4(p.n.x, p.n.y, p.n.z, p.sw);
    g[2 * P + o] = make_float4(p.V.x, p.V.y, p.V.z, 0.0f);
    g[3 * P + o] = make_float4(p.a.x, p.a.y, p.a.z, 0.0f);
}

// One workgroup: every listed parameter, then the step counter and the loss
// history (the loss is that of the parameters before this step).
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ par, const float* __restrict__ grad,
                                              const float* __restrict__ base, float* __restrict__ theta,
                                              float* __restrict__ m, float* __restrict__ v,
                                              const int32_t* __restrict__ idx, const float* __restrict__ lower,
                                              AdamConsts a, int32_t* step, const double* loss, double* hist) {
    const int t = *step + 1;
    const float c1 = (float)(1.0 - pow64((double)a.beta1, (double)t));
    const float c2 = (float)(1.0 - pow64((double)a.beta2, (double)t));
    for (int k = threadIdx.x; k < a.n; k += blockDim.x) {
        const int i = idx[k];
        const float b0 = base[i];
        const float g = grad[i] * b0;  // d loss / d theta for p = base * (1 + theta)
        const float mk = a.beta1 * m[k] + (1.0f - a.beta1) * g;
        const float vk = a.beta2 * v[k] + (1.0f - a.beta2) * (g * g);
        m[k] = mk, v[k] = vk;
        float th = theta[k] - a.lr * (mk / c1) / (__builtin_sqrtf(vk / c2) + a.eps);
        float pv = b0 * (1.0f + th);
        if (pv < lower[k]) {
            pv = lower[k];
            if (b0 != 0.0f) th = pv / b0 - 1.0f;
        }
        theta[k] = th, par[i] = pv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *step = t;
        if (t <= a.hist_cap) hist[t - 1] = *loss;
    }
}

void launch_diff(hipStream_t st, bool backward, const DevScene& s, const ViewConsts& v, const DiffBuffers& b) {
    const int grid = diff_blocks(v);
    if (backward) hipLaunchKernelGGL(k_diff<true>, dim3(grid), dim3(256), 0, st, s, v, b);
    else hipLaunchKernelGGL(k_diff<false>, dim3(grid), dim3(256), 0, st, s, v, b);
}

int32_t diff_blocks(const ViewConsts& v) {
    const int n_tiles = v.tiles_x * v.tiles_y;
    return ((n_tiles + 7) / 8) * 8;
}

void launch_diff_views(hipStream_t st, bool backward, const DevScene& s, const DiffView* views, int32_t n_views,
                       int32_t total_blocks) {
    if (n_views <= 0 || total_blocks <= 0) return;
    // (a fixed grid of 2048 workgroups, 8 per CU, strides over the views' tiles)
    const int grid = total_blocks < 2048 ? total_blocks : 2048;
    if (backward)
        hipLaunchKernelGGL(k_diff_views<true>, dim3(grid), dim3(256), 0, st, s, views, n_views, total_blocks);
    else hipLaunchKernelGGL(k_diff_views<false>, dim3(grid), dim3(256), 0, st, s, views, n_views, total_blocks);
}

void launch_diff_store_gbuffer(hipStream_t st, const DevScene& s, const ViewConsts& v, const DiffBuffers& b,
                               float4* gbuf) {
    const size_t P = (size_t)v.W * v.H;
    if (P) hipLaunchKernelGGL(k_diff_gbuffer, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, s, v, b, gbuf);
}

void launch_adam(hipStream_t st, float* par, const float* grad, const float* base, float* theta, float* m, float* v,
                 const int32_t* idx, const float* lower, AdamConsts a, int32_t* step, const double* loss,
                 double* hist) {
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, st, par, grad, base, theta, m, v, idx, lower, a, step, loss,
                       hist);
}

}  // namespace prh
