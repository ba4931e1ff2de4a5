// kernels.hip — libprh's gfx950 kernels: vertex transform, triangle setup /
// clip / bin, LDS-binned visibility and shadow rasters, deferred shading with
// shadow maps, SSAO and sRGB fused.  All arithmetic that decides coverage,
// depth, primitive order or colour follows polyred's CPU renderer bit for bit
// (render/raster.go, render/clipping.go, render/shadow.go,
// shader/blinn_cpu.go, material/ao.go, shader/gamma.go) through gomath.hpp.
#include "kernels.hpp"

#include <functional>
#include <map>
#include <vector>

using namespace gm;

namespace prh {

namespace {

// 24-bit multiplies as the one full-rate instruction.  __umul24/__mul24 in a
// loop whose other operand is loop-invariant compile to a mask hoisted out of
// the loop and a quarter-rate v_mul_lo_u32 inside it.
__device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ int32_t mul_i24(int32_t a, int32_t b) {
    int32_t r;
    asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ V4 ld(float4 a) { return v4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// Go's math.Min/Max folds (NaN- and Inf-aware, -0 below +0) reduce to plain
// min/max when every input is finite; the sign of a zero extreme never
// changes a result here (they feed comparisons and Round), so finite inputs
// take the fast path and anything else the exact fold.
__device__ __forceinline__ bool fin(float a) { return __builtin_isfinite(a); }
__device__ __forceinline__ float gmin3(float a, float b, float c) {
    if (fin(a) && fin(b) && fin(c)) return fminf(fminf(a, b), c);
    return min2(min2(min2(FLT_MAX, a), b), c);
}
__device__ __forceinline__ float gmax3(float a, float b, float c) {
    if (fin(a) && fin(b) && fin(c)) return fmaxf(fmaxf(a, b), c);
    return max2(max2(max2(-FLT_MAX, a), b), c);
}

// Pixel box of a triangle, render/raster.go:473-477 (and shadow.go:187-191):
// int(Round(min)-1) .. int(Round(max)+1) from the AABB of the three vertices,
// then restricted to the buffer (FragmentBuffer.In).  Returns false if empty.
//
// For a small, well-shaped triangle the box is then narrowed to the pixels
// whose centre lies within 0.24 of the AABB; the pixels dropped all fail the
// inside test (no Barycoord >= -1e-7), so the result is the same.  Why: for a
// centre p at least d = 0.23 outside the AABB along x (y alike), the exact
// barycentrics l satisfy p.x - min.x = sum l_i (x_i - min.x) with at most two
// l_i < 0, so some l_i <= -d / (2 W), W the AABB's larger side.  Over the box
// (within W + 2 of the vertices) |l_i| <= 2 (W + 2) / (s W) when
// |Sabc| >= s W^2, and Barycoord's float32 evaluation (the differences, the
// fma64 areas, the division) stays within about 300 ulp * (1 + 2 / W) of l
// for s = 1/4: under 2e-5 (W <= 64) or 4e-5 / W, far from closing the gap
// from -0.115 / W to -1e-7.  Slivers, larger or degenerate triangles and
// non-finite vertices keep the box.
__device__ bool pixel_box(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz,
                          int W, int H, int& x0, int& y0, int& x1, int& y1) {
    float mnx = gmin3(ax, bx, cx), mny = gmin3(ay, by, cy);
    float mxx = gmax3(ax, bx, cx), mxy = gmax3(ay, by, cy);
    (void)az, (void)bz, (void)cz;
    int64_t xmin = f2i(round32(mnx) - 1), xmax = f2i(round32(mxx) + 1);
    int64_t ymin = f2i(round32(mny) - 1), ymax = f2i(round32(mxy) + 1);
    const float wmax = fmaxf(mxx - mnx, mxy - mny);
    if (wmax <= 64.0f && wmax >= 0x1p-20f) {  // false for NaN / Inf
        const float sabc = tri_setup(ax, ay, bx, by, cx, cy).sabc;
        if (fabsf(sabc) >= 0.25f * wmax * wmax) {
            // centres x + 0.5 in [min - 0.24, max + 0.24]
            xmin = max(xmin, (int64_t)ceilf(mnx - 0.74f)), xmax = min(xmax, (int64_t)floorf(mxx - 0.26f));
            ymin = max(ymin, (int64_t)ceilf(mny - 0.74f)), ymax = min(ymax, (int64_t)floorf(mxy - 0.26f));
        }
    }
    if (xmin < 0) xmin = 0;
    if (ymin < 0) ymin = 0;
    if (xmax > W - 1) xmax = W - 1;
    if (ymax > H - 1) ymax = H - 1;
    if (xmin > xmax || ymin > ymax) return false;
    x0 = (int)xmin, y0 = (int)ymin, x1 = (int)xmax, y1 = (int)ymax;
    return true;
}

struct Box3 {
    float mnx, mny, mnz, mxx, mxy, mxz;
};
__device__ Box3 box_of(V4 a, V4 b, V4 c) {  // primitive.NewAABB, box.go:15-25
    Box3 r;
    r.mnx = gmin3(a.x, b.x, c.x), r.mny = gmin3(a.y, b.y, c.y), r.mnz = gmin3(a.z, b.z, c.z);
    r.mxx = gmax3(a.x, b.x, c.x), r.mxy = gmax3(a.y, b.y, c.y), r.mxz = gmax3(a.z, b.z, c.z);
    return r;
}
// AABB.Intersect with the viewport box Min (0,0,-1), Max (w,h,1); the z test
// compares against the viewport's Max.Y (geometry/primitive/box.go:38).
__device__ bool vp_intersect(const Box3& t, float w, float h) {
    if (fin(t.mnx) && fin(t.mny) && fin(t.mnz) && fin(t.mxx) && fin(t.mxy) && fin(t.mxz))
        return fmaxf(0.0f, t.mnx) <= fminf(w, t.mxx) && fmaxf(0.0f, t.mny) <= fminf(h, t.mxy) &&
               fmaxf(-1.0f, t.mnz) <= fminf(h, t.mxz);
    float minX = max2(0.0f, t.mnx), minY = max2(0.0f, t.mny), minZ = max2(-1.0f, t.mnz);
    float maxX = min2(w, t.mxx), maxY = min2(h, t.mxy), maxZ = min2(h, t.mxz);
    return minX <= maxX && minY <= maxY && minZ <= maxZ;
}
// (bitwise & and |: every term is a pure comparison, so evaluating them all
// gives the same answer without a branch per term)
__device__ __forceinline__ bool less_eq(float v1, float v2) {  // box.go:62-64
    const float eps = 1e-7f;
    return approx_eq(v1, v2, eps) | ((v1 < v2) & (abs32(v1 - v2) > eps));
}
__device__ bool vp_contains(V4 p, float w, float h) {  // box.go:61-75 for one vertex
    return less_eq(0.0f, p.x) & less_eq(0.0f, p.y) & less_eq(-1.0f, p.z) & less_eq(p.x, w) & less_eq(p.y, h) &
           less_eq(p.z, 1.0f);
}

// Adds one to counter[key] for every active lane with a single atomic per
// distinct key in the wave (neighbouring triangles share tiles, and a shadow
// map concentrates a scene in few tiles, so per-lane atomics collide).
__device__ __forceinline__ void wave_count(int32_t* counter, int key, bool active) {
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        if (__lane_id() == leader) atomicAdd(&counter[lkey], __popcll(same));
        pending &= ~same;
    }
}
// counter[key]++ for every active lane, one atomic per distinct key; returns
// the lane's pre-increment value.
__device__ __forceinline__ int wave_alloc(int32_t* counter, int key, bool active) {
    uint64_t pending = __ballot(active);
    int pos = 0;
    const int lane = __lane_id();
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(&counter[lkey], __popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) pos = base + __popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return pos;
}
// Counts a primitive into every tile its pixel box touches (wave-uniform call).
__device__ __forceinline__ void count_tiles(int32_t* tile_count, int base, int tiles_x, int shift, bool active, int x0,
                                            int y0, int x1, int y1) {
    const int tx0 = x0 >> shift, ty0 = y0 >> shift, ntx = active ? (x1 >> shift) - tx0 + 1 : 0;
    const int n = active ? ntx * ((y1 >> shift) - ty0 + 1) : 0;
    for (int j = 0; __ballot(j < n) != 0; j++) {
        const bool act = j < n;
        const int key = act ? base + (ty0 + j / ntx) * tiles_x + tx0 + j % ntx : 0;
        wave_count(tile_count, key, act);
    }
}

constexpr uint32_t PRIM_EMPTY = 0xffffffffu;  // bbox word of an unused primitive slot

// A primitive slot is its 48-byte record plus its pixel box in a separate
// 8-byte array, which is all k_bin reads; an empty slot only marks its box.
__device__ __forceinline__ void store_empty(uint2* pbox, size_t slot) {
    pbox[slot] = make_uint2(PRIM_EMPTY, PRIM_EMPTY);
}
__device__ __forceinline__ void store_prim(Prim* prims, uint2* pbox, size_t slot, V4 a, V4 b, V4 c, uint32_t key,
                                           int x0, int y0, int x1, int y1) {
    Prim p;
    p.p0 = make_float4(a.x, a.y, a.z, b.x);
    p.p1 = make_float4(b.y, b.z, c.x, c.y);
    p.p2 = make_float4(c.z, bitsf(key), bitsf(pack_bbox(x0, y0)), bitsf(pack_bbox(x1, y1)));
    prims[slot] = p;
    pbox[slot] = make_uint2(pack_bbox(x0, y0), pack_bbox(x1, y1));
}

// The six screen-space planes of Sutherland-Hodgman, clipping.go:31-65
// (the clipper itself is clip_wave below).
struct Plane {
    V4 pos, nor;
};
__device__ __forceinline__ bool in_front(const Plane& p, V4 v) { return dot(sub(v, p.pos), p.nor) > 0; }
__device__ __forceinline__ V4 isect(const Plane& p, V4 v0, V4 v1) {
    V4 u = sub(v1, v0), w = sub(v0, p.pos);
    float d = dot(p.nor, u);
    float n = -dot(p.nor, w);
    float s = n / d;
    return add(v0, scale(u, s, s, s, s));
}
__device__ __forceinline__ uint32_t u8x4(uint8_t r, uint8_t g, uint8_t b, uint8_t a) {
    return (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24);
}
__device__ __forceinline__ float chan(uint32_t c, int k) { return (float)((c >> (8 * k)) & 0xff); }

}  // namespace

// A kernel body's workgroup: its index and the grid's size in x and y --
// blockIdx/gridDim for a single-view launch, the view-local values in a
// wide launch.
struct VB {
    int bx, by, gx, gy;
};
// The view of workgroup b in a wide launch: the last i with start[i] <= b.
__device__ __forceinline__ int wide_view(const int32_t* __restrict__ start, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
// The frame's start: counters to zero and the visibility buffer empty.
struct k_frame_clear_args {
    uint32_t* cnt;
    int32_t n_cnt;
    uint64_t* vis;
    size_t n_vis;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void frame_clear_part(const FrameClear& c, size_t i0, size_t stride) {
    if (i0 < (size_t)c.n_cnt) c.cnt[i0] = 0u;
    for (size_t i = i0; i < c.n_vis; i += stride) c.vis[i] = VIS_EMPTY;
}
__device__ __forceinline__ void k_frame_clear_body(const k_frame_clear_args& a_, const VB vb) {
    frame_clear_part(FrameClear{a_.cnt, a_.n_cnt, a_.vis, a_.n_vis}, vb.bx * (size_t)blockDim.x + threadIdx.x,
                     (size_t)vb.gx * blockDim.x);
}
__global__ __launch_bounds__(256) void k_frame_clear(k_frame_clear_args a) {
    k_frame_clear_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_frame_clear_wide(const k_frame_clear_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_frame_clear_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_frame_clear_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------- vertex transform
// render/raster.go:381-410 per vertex: Pos = trans.MulV(p) with
// trans = Proj.View.Model, recipw = -1/W (perspective), Pos.Apply(Viewport).Pos();
// Nor = n.Apply(Normal); and drawClipped's back transform
// m = Pos.Apply(ViewportInv).Apply(ProjInv).Apply(ViewInv) (raster.go:467-469).
struct k_vertex_main_args {
    DevScene s;
    const M4* trans;
    const M4* normal;
    ViewConsts vc;
    float4* tv_scr;
    float4* tv_nor;
    float4* tv_wp;
    uint32_t* tv_col;
    float2* tv_uv;
    FrameClear clr;  // the frame's start, done by this pass
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_vertex_main_body(const k_vertex_main_args& a_, const VB vb) {
    frame_clear_part(a_.clr, vb.bx * (size_t)blockDim.x + threadIdx.x, (size_t)vb.gx * blockDim.x);
    const DevScene& s = a_.s;
    const M4* __restrict__ trans = a_.trans;
    const M4* __restrict__ normal = a_.normal;
    const ViewConsts& vc = a_.vc;
    float4* __restrict__ tv_scr = a_.tv_scr;
    float4* __restrict__ tv_nor = a_.tv_nor;
    float4* __restrict__ tv_wp = a_.tv_wp;
    uint32_t* __restrict__ tv_col = a_.tv_col;
    float2* __restrict__ tv_uv = a_.tv_uv;

    int i = vb.bx * blockDim.x + threadIdx.x;
    if (i >= s.n_tverts) return;
    int o = s.vtx_obj[i];
    int mv = s.mesh_vtx_off[s.obj_mesh[o]] + (i - s.obj_vtx_start[o]);
    V4 p = ld(s.vpos[mv]), n = ld(s.vnor[mv]);
    V4 clip = mulv(trans[o], p);
    float rw = vc.perspect ? -1.0f / clip.w : 1.0f;
    V4 sp = pos(apply(clip, vc.vp));
    V4 nn = apply(n, normal[o]);
    V4 m = apply(apply(apply(sp, vc.vpinv), vc.pinv), vc.vinv);
    tv_scr[i] = make_float4(sp.x, sp.y, sp.z, rw);
    tv_nor[i] = make_float4(nn.x, nn.y, nn.z, 0.0f);
    tv_wp[i] = make_float4(m.x, m.y, m.z, 0.0f);
    // the vertex colour and UV by transformed vertex, for the shading
    tv_col[i] = s.vcol[mv];
    if (tv_uv) tv_uv[i] = s.vuv[mv];
}
__global__ __launch_bounds__(256) void k_vertex_main(k_vertex_main_args a) {
    k_vertex_main_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_vertex_main_wide(const k_vertex_main_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_vertex_main_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_vertex_main_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// render/shadow.go:155-173: light-space screen position of every vertex for
// every shadow slot (a thread per vertex, which reads it once for all the
// slots).
struct k_vertex_shadow_args {
    DevScene s;
    const M4* strans;
    ShadowConsts sc;
    float3* ts;  // [slot][vertex] light-space screen x y z
    FrameClear clr;  // the shadow passes' counters, cleared by this pass
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_vertex_shadow_body(const k_vertex_shadow_args& a_, const VB vb) {
    frame_clear_part(a_.clr, vb.bx * (size_t)blockDim.x + threadIdx.x, (size_t)vb.gx * blockDim.x);
    const DevScene& s = a_.s;
    const M4* __restrict__ strans = a_.strans;
    const ShadowConsts& sc = a_.sc;
    float3* __restrict__ ts = a_.ts;

    const int i = vb.bx * blockDim.x + threadIdx.x;
    if (i >= s.n_tverts) return;
    const int o = s.vtx_obj[i];
    const int mv = s.mesh_vtx_off[s.obj_mesh[o]] + (i - s.obj_vtx_start[o]);
    const V4 v = ld(s.vpos[mv]);
    for (int slot = 0; slot < sc.n_slots; slot++) {
        const V4 q = pos(apply(mulv(strans[(size_t)slot * s.n_objects + o], v), sc.vp));
        ts[(size_t)slot * s.n_tverts + i] = make_float3(q.x, q.y, q.z);
    }
}
__global__ __launch_bounds__(256) void k_vertex_shadow(k_vertex_shadow_args a) {
    k_vertex_shadow_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_vertex_shadow_wide(const k_vertex_shadow_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_vertex_shadow_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_vertex_shadow_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------- triangle setup
// render/raster.go:257-269 (IsValid) and draw, raster.go:380-444: back-face
// cull, viewport AABB test, then either one primitive in the triangle's own
// slot (all vertices inside) or a place in the clipper's queue.  Each
// primitive is counted into every 16x16 tile its pixel box touches.
// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ void block_scan_256(int v, int32_t* start /*[257]*/, int32_t* wsum /*[4]*/) {
    int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        int u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < wv; k++) before += wsum[k];
    start[tid] = before + incl - v;
    if (tid == 255) start[256] = before + incl;
    __syncthreads();
}

// Inclusive scan of one int per lane over the wave (sum or max, values >= 0)
// in DPP steps: Hillis-Steele over each row of 16 lanes (row_shr 1, 2, 4,
// 8; a lane past its row's start reads 0, the identity of both), then the
// last lane of row 0 and of row 2 into rows 1 and 3 (row_bcast:15), and the
// last lane of row 1 into rows 2 and 3 (row_bcast:31).
template <bool MAX>
__device__ __forceinline__ int wave_scan_incl(int v) {
    auto op = [](int a, int b) { return MAX ? max(a, b) : a + b; };
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
    return v;
}
// LDS written by some lanes of a wave and read by others: visible past a
// wave-level fence (no workgroup barrier needed)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Each lane's c items spread over the wave's lanes, 64 a round: a wave
// scan places them, and per round each lane whose items start in the round
// marks its start slot (LDS, 64 a wave); a max-scan of the marks (and of
// the lane running into the round) names each slot's owner.  body(o, l,
// valid) runs on every lane each round (so that it may gather from other
// lanes), for item l of lane o where valid.
template <class F>
__device__ __forceinline__ void wave_spread(int c, int32_t* wmark /*[64]*/, F&& body) {
    const int lane = threadIdx.x & 63;
    const int incl = wave_scan_incl<false>(c);
    const int s0 = incl - c;  // this lane's first item
    const int total = __builtin_amdgcn_readlane(incl, 63);
    int carry = 0;  // 1 + the lane owning the slot before the round
    for (int base = 0; base < total; base += 64) {
        wave_lds_sync();  // (the previous round's reads, this round's writes)
        wmark[lane] = 0;
        wave_lds_sync();
        if (c > 0 && s0 >= base && s0 < base + 64) wmark[s0 - base] = lane + 1;
        wave_lds_sync();
        int m = wmark[lane];
        if (lane == 0) m = max(m, carry);
        m = wave_scan_incl<true>(m);
        carry = __builtin_amdgcn_readlane(m, 63);
        const int idx = base + lane, o = max(m - 1, 0);
        const int so = __builtin_amdgcn_ds_bpermute(o * 4, s0);
        body(o, idx - so, idx < total && m > 0);
    }
}

// Small triangles drawn by the setup workgroups themselves: each wave's
// candidate pixel centres (its lanes' pixel boxes) are spread over its 64
// lanes (wave_spread), each centre tested as the rasters test it (the same
// Barycoord, inside test and depth, raster.go:487-496) and handed to `put`
// with the lane of its triangle, whose record is staged in LDS.
struct __align__(16) DPrim {
    float ax, ay, bx, by, abx, aby, acx, acy, bcx, bcy, sabc;  // TriSetup
    float az, bz, cz;
    int32_t xy0;  // x0 | y0 << 16
    uint32_t wq;  // box width w | ceil(65536 / w) << 9
};
template <class F>
__device__ __forceinline__ void draw_small_tris(bool mine, V4 p1, V4 p2, V4 p3, int x0, int y0, int x1, int y1,
                                                DPrim* sp /*[256]*/, int32_t* mark /*[256]*/, F&& put) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    DPrim* wsp = sp + 64 * wv;
    int c = 0;
    if (mine) {
        const TriSetup t = tri_setup(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
        DPrim& d = wsp[lane];
        d.ax = t.ax, d.ay = t.ay, d.bx = t.bx, d.by = t.by, d.abx = t.abx, d.aby = t.aby;
        d.acx = t.acx, d.acy = t.acy, d.bcx = t.bcx, d.bcy = t.bcy, d.sabc = t.sabc;
        d.az = p1.z, d.bz = p2.z, d.cz = p3.z;
        const uint32_t w = (uint32_t)(x1 - x0 + 1);
        d.xy0 = x0 | (y0 << 16), d.wq = w | ((65536u + w - 1u) / w) << 9;
        c = (x1 - x0 + 1) * (y1 - y0 + 1);
    }
    wave_spread(c, mark + 64 * wv, [&](int o, int l, bool valid) {
        if (!valid) return;
        const DPrim d = wsp[o];
        const int w = (int)(d.wq & 511u);
        // l / w by multiply-shift: exact while l * w < 65536, which l < area <= 256
        // and w <= area keep (the thresholds are at most 256)
        const int q = (int)(mul_u24((uint32_t)l, d.wq >> 9) >> 16);
        const int px = (d.xy0 & 0xffff) + (l - q * w), py = (d.xy0 >> 16) + q;
        TriSetup t;
        t.ax = d.ax, t.ay = d.ay, t.bx = d.bx, t.by = d.by, t.abx = d.abx, t.aby = d.aby;
        t.acx = d.acx, t.acy = d.acy, t.bcx = d.bcx, t.bcy = d.bcy, t.sabc = d.sabc;
        float z;
        if (cover_z(t, (float)px + 0.5f, (float)py + 0.5f, d.az, d.bz, d.cz, z)) put(px, py, z, o);
    });
}

// An unclipped triangle whose pixel box holds at most `draw_small` pixel
// centres (0: none; prh_tuning.main_direct_max) is drawn by the setup itself
// outside the exact-order pass: each
// centre is tested as k_raster<VIS> tests it and merged into the visibility
// buffer with a global 64-bit atomic min of the same key (order-independent),
// and the triangle is neither stored nor binned.

struct k_setup_main_args {
    DevScene s;
    ViewConsts vc;
    const float4* tv_scr;
    Prim* prims;
    uint2* pbox;
    ClipJob* clip_list;
    int32_t* tile_count;
    int32_t* blk_prims;
    Counters* cnt;
    Capacities cap;
    unsigned long long* vis;
    int32_t draw_small;
    const int4* tri_rec;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_setup_main_body(const k_setup_main_args& a_, const VB vb) {
    const DevScene& s = a_.s;
    const ViewConsts& vc = a_.vc;
    const float4* __restrict__ tv_scr = a_.tv_scr;
    Prim* __restrict__ prims = a_.prims;
    uint2* __restrict__ pbox = a_.pbox;
    ClipJob* __restrict__ clip_list = a_.clip_list;
    int32_t* __restrict__ tile_count = a_.tile_count;
    int32_t* __restrict__ blk_prims = a_.blk_prims;
    Counters* cnt = a_.cnt;
    const Capacities& cap = a_.cap;
    unsigned long long* __restrict__ vis = a_.vis;
    const int32_t draw_small = a_.draw_small;
    const int4* __restrict__ tri_rec = a_.tri_rec;

    const int g = vb.bx * blockDim.x + threadIdx.x;
    bool direct = false, doclip = false, drawn = false;
    ClipJob job{g, 0, 0, 0, 0, {0, 0, 0}};
    V4 p1 = v4(0, 0, 0, 1), p2 = p1, p3 = p1;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    const float w = vc.vpbox_max_x, h = vc.vpbox_max_y;
    if (g < s.n_tris) {
        const int4 rec = tri_rec[g];  // (x < 0: not Triangle.IsValid)
        if (rec.x >= 0) {
            const int i0 = rec.x, i1 = rec.y, i2 = rec.z;
            const float4 A = tv_scr[i0], B = tv_scr[i1], Cc = tv_scr[i2];
            p1 = v4(A.x, A.y, A.z, 1), p2 = v4(B.x, B.y, B.z, 1), p3 = v4(Cc.x, Cc.y, Cc.z, 1);
            if (!(cross(sub(p2, p1), sub(p3, p1)).z < 0)) {  // cull.go:26-28
                Box3 tb = box_of(p1, p2, p3);
                if (vp_intersect(tb, w, h)) {
                    if (vp_contains(p1, w, h) & vp_contains(p2, w, h) & vp_contains(p3, w, h))
                        direct = pixel_box(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, vc.W, vc.H, x0, y0,
                                           x1, y1);
                    else {
                        // (the offset from the transformed vertices to the mesh's)
                        const int o = s.tri_obj[g];
                        doclip = true, job.i0 = i0, job.i1 = i1, job.i2 = i2;
                        job.dv = s.mesh_vtx_off[s.obj_mesh[o]] - s.obj_vtx_start[o];
                    }
                }
            }
        }
        const uint32_t key = (uint32_t)g * 8u + SUB_DIRECT;
        drawn = direct && draw_small > 0 && (x1 - x0 + 1) * (y1 - y0 + 1) <= draw_small;
        if (direct && !drawn) store_prim(prims, pbox, g, p1, p2, p3, key, x0, y0, x1, y1);
        else store_empty(pbox, g);
    }
    {
        __shared__ DPrim sp[256];
        __shared__ int32_t mark[256];
        const uint32_t key0 = (uint32_t)(g - (int)(threadIdx.x & 63)) * 8u + SUB_DIRECT;  // the wave's first triangle
        bool nan = false;
        draw_small_tris(drawn, p1, p2, p3, x0, y0, x1, y1, sp, mark, [&](int px, int py, float z, int o) {
            if (z != z) nan = true;
            else if ((unsigned)px < (unsigned)vc.W && (unsigned)py < (unsigned)vc.H)  // (inside by construction)
                atomicMin(&vis[(size_t)py * vc.W + px],
                          ((unsigned long long)depth_key_hi(z) << 32) | (key0 + 8u * (uint32_t)o));
        });
        if (nan) atomicOr(&cnt->nan_flag, 1);
    }
    count_tiles(tile_count, 0, vc.rtiles_x, vc.rt_shift, direct && !drawn, x0, y0, x1, y1);
    const int pos = wave_alloc(&cnt->n_clipped, 0, doclip);
    if (doclip) {
        if (pos < cap.clip_tris) clip_list[pos] = job;
        else atomicOr(&cnt->overflow, 1);
    }
    // the primitive count of the frame statistics, summed by k_scan (one
    // global counter hit by every workgroup would serialise them)
    const int nd = __syncthreads_count(direct);
    if (threadIdx.x == 0) blk_prims[vb.bx] = nd;
}
__global__ __launch_bounds__(256) void k_setup_main(k_setup_main_args a) {
    k_setup_main_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_setup_main_wide(const k_setup_main_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_setup_main_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_setup_main_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// render/shadow.go:152-176 (drawDepth up to the pixel loop) for every
// (triangle, shadow slot): back-face cull, cullViewFrustum, pixel box.
// A shadow-map triangle whose pixel box holds at most sc.direct_max texel
// centres (prh_tuning.shadow_direct_max) is drawn here: each centre is
// tested and merged into the map with a
// global atomic max, exactly as k_raster<SHADOW> would (the same Barycoord,
// depth and max over ordered bits; order-independent), and it is neither
// stored nor binned.  Most triangles of a map fitted to the whole frustum
// cover a few texels; past a few dozen the binned raster is no slower.

struct k_setup_shadow_args {
    DevScene s;
    ShadowConsts sc;
    const int4* tri_rec;
    const float3* ts;
    Prim* prims;
    uint2* pbox;
    int32_t* tile_count;
    Counters* cnt;
    Capacities cap;
    uint32_t* smaps;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_setup_shadow_body(const k_setup_shadow_args& a_, const VB vb) {
    const DevScene& s = a_.s;
    const ShadowConsts& sc = a_.sc;
    const float3* __restrict__ ts = a_.ts;
    Prim* __restrict__ prims = a_.prims;
    uint2* __restrict__ pbox = a_.pbox;
    int32_t* __restrict__ tile_count = a_.tile_count;
    Counters* cnt = a_.cnt;
    uint32_t* __restrict__ smaps = a_.smaps;

    const int g = vb.bx * blockDim.x + threadIdx.x;
    const int slot = vb.by;
    bool emit = false;
    V4 p1 = v4(0, 0, 0, 1), p2 = p1, p3 = p1;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    if (g < s.n_tris) {
        const int4 rec = a_.tri_rec[g];  // (x < 0: not Triangle.IsValid)
        if (rec.x >= 0) {
            const float3* T = ts + (size_t)slot * s.n_tverts;
            const float3 A = T[rec.x], B = T[rec.y], C = T[rec.z];
            p1 = v4(A.x, A.y, A.z, 1), p2 = v4(B.x, B.y, B.z, 1), p3 = v4(C.x, C.y, C.z, 1);
            if (!(cross(sub(p2, p1), sub(p3, p1)).z < 0)) {
                Box3 tb = box_of(p1, p2, p3);
                if (vp_intersect(tb, sc.vpbox_max_x, sc.vpbox_max_y))
                    emit = pixel_box(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, sc.SW, sc.SH, x0, y0, x1, y1);
            }
        }
    }
    const bool small = emit && (x1 - x0 + 1) * (y1 - y0 + 1) <= sc.direct_max;
    if (small) emit = false;
    {
        __shared__ DPrim sp[256];
        __shared__ int32_t mark[256];
        uint32_t* map = smaps + (size_t)slot * sc.SW * sc.SH;
        bool nan = false;
        draw_small_tris(small, p1, p2, p3, x0, y0, x1, y1, sp, mark, [&](int px, int py, float z, int) {
            if (z != z) nan = true;
            else if (z > 0.0f && (unsigned)px < (unsigned)sc.SW && (unsigned)py < (unsigned)sc.SH)  // (inside by construction)
                atomicMax(&map[(size_t)py * sc.SW + px], depth_ord(z));  // the map starts at 0
        });
        if (nan) atomicOr(&cnt->shadow_nan, 1);
    }
    if (g < s.n_tris) {
        const size_t ps = (size_t)slot * s.n_tris + g;
        if (emit) store_prim(prims, pbox, ps, p1, p2, p3, (uint32_t)slot, x0, y0, x1, y1);
        else store_empty(pbox, ps);
    }
    count_tiles(tile_count, slot * sc.tiles_per_map, sc.tiles_x, sc.rt_shift, emit, x0, y0, x1, y1);
}
__global__ __launch_bounds__(256) void k_setup_shadow(k_setup_shadow_args a) {
    k_setup_shadow_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_setup_shadow_wide(const k_setup_shadow_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_setup_shadow_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_setup_shadow_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------------ tile scan
// Exclusive prefix sums of the per-tile primitive counts and of the per-tile
// work items (ceil(count / ITEM_PRIMS)): per-thread runs of consecutive
// tiles, wave scans, then the 16 wave totals.  Up to SCAN_WHOLE_MAX tiles
// (a 1080p view's 8160) one workgroup scans them all, SCAN_WHOLE_MAX / 1024
// a thread, in one launch; beyond, k_scan_sums reduces each block of SCAN_BLOCK tiles
// and every workgroup of k_scan_apply adds the sums of the blocks before its
// own and scans its block, 4 tiles a thread.
constexpr int SCAN_BLOCK = 4096;
constexpr int SCAN_WHOLE_MAX = 8192;

__device__ __forceinline__ int wave_incl_scan(int v, int) { return wave_scan_incl<false>(v); }  // (v >= 0)
__device__ __forceinline__ int wave_sum(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int item_count(int v, int ip) { return (v + ip - 1) / ip; }

struct k_scan_sums_args {
    const int32_t* counts;
    int32_t n;
    int32_t* bsum;
    int32_t item_prims;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_scan_sums_body(const k_scan_sums_args& a_, const VB vb) {
    const int32_t item_prims = a_.item_prims;
    const int32_t* __restrict__ counts = a_.counts;
    const int32_t n = a_.n;
    int32_t* __restrict__ bsum = a_.bsum;

    __shared__ int32_t ws[16], wc[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i0 = vb.bx * SCAN_BLOCK + 4 * tid;
    int sv = 0, sc = 0;
    for (int k = 0; k < 4; k++) {
        const int v = i0 + k < n ? counts[i0 + k] : 0;
        sv += v;
        sc += item_count(v, item_prims);
    }
    sv = wave_sum(sv), sc = wave_sum(sc);
    if (lane == 0) ws[wv] = sv, wc[wv] = sc;
    __syncthreads();
    if (tid == 0) {
        int a = 0, c = 0;
        for (int k = 0; k < 16; k++) a += ws[k], c += wc[k];
        bsum[2 * vb.bx] = a;
        bsum[2 * vb.bx + 1] = c;
    }
}
__global__ __launch_bounds__(1024) void k_scan_sums(k_scan_sums_args a) {
    k_scan_sums_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(1024) void k_scan_sums_wide(const k_scan_sums_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_scan_sums_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_scan_sums_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// Also lists each work item's tile (item_tile, up to cap_items entries) and
// leaves the counts at zero for the next pass.
struct k_scan_apply_args {
    int32_t* counts;
    int32_t n;
    const int32_t* bsum;
    int32_t* offsets;
    int32_t* chunk_off;
    int32_t* cursor;
    int32_t* item_tile;
    int32_t cap_items;
    const int32_t* blk_prims;
    int32_t n_blk;
    Counters* cnt;
    int32_t item_prims;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_scan_apply_body(const k_scan_apply_args& a_, const VB vb) {
    constexpr int NT = 1024, NW = NT / 64, PER = SCAN_WHOLE_MAX / NT;
    const int32_t item_prims = a_.item_prims;
    int32_t* __restrict__ counts = a_.counts;
    const int32_t n = a_.n;
    const int32_t* __restrict__ bsum = a_.bsum;
    int32_t* __restrict__ offsets = a_.offsets;
    int32_t* __restrict__ chunk_off = a_.chunk_off;
    int32_t* __restrict__ cursor = a_.cursor;
    int32_t* __restrict__ item_tile = a_.item_tile;
    const int32_t cap_items = a_.cap_items;
    const int32_t* __restrict__ blk_prims = a_.blk_prims;
    const int32_t n_blk = a_.n_blk;
    Counters* cnt = a_.cnt;

    __shared__ int32_t ws[NW], wc[NW];
    __shared__ int32_t carry_s, carry_c;
    __shared__ int32_t s_t[SCAN_WHOLE_MAX + SCAN_WHOLE_MAX / 32];  // whole scan: tile i at i + i / 32
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool whole = vb.gx == 1;
    // the frame statistics' per-workgroup primitive counts, loaded at once
    // by the workgroup that sums them (the last)
    int np = 0;
    if (whole || vb.bx == vb.gx - 1) {
        int pv[8];  // (every load issued before the first is waited for)
#pragma unroll
        for (int j = 0; j < 8; j++) pv[j] = tid + j * NT < n_blk ? blk_prims[tid + j * NT] : 0;
#pragma unroll
        for (int j = 0; j < 8; j++) np += pv[j];
        for (int k = tid + 8 * NT; k < n_blk; k += NT) np += blk_prims[k];
    }
    int run = 0, runc = 0;
    if (whole) {
        // One workgroup for every tile: thread t takes tiles PER*t ..
        // PER*t + PER-1, read and written through LDS so that the global
        // accesses stay coalesced (the padded slots keep both LDS patterns
        // free of bank conflicts).
        {
            int c[PER];  // (every load issued before the first is waited for)
#pragma unroll
            for (int k = 0; k < PER; k++) c[k] = k * NT + tid < n ? counts[k * NT + tid] : 0;
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = k * NT + tid;
                if (i < n) s_t[i + (i >> 5)] = c[k], counts[i] = 0;
            }
        }
        __syncthreads();
        int v[PER], ch[PER];
        int local = 0, localc = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = PER * tid + k;
            v[k] = i < n ? s_t[i + (i >> 5)] : 0;
            ch[k] = item_count(v[k], item_prims);
            local += v[k], localc += ch[k];
        }
        const int incl = wave_incl_scan(local, lane), inclc = wave_incl_scan(localc, lane);
        if (lane == 63) ws[wv] = incl, wc[wv] = inclc;
        __syncthreads();  // (and every thread has read its tiles)
        run = incl - local, runc = inclc - localc;
        for (int k = 0; k < wv; k++) run += ws[k], runc += wc[k];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = PER * tid + k;
            if (i < n) s_t[i + (i >> 5)] = run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = k * NT + tid;
            if (i < n) {
                const int o = s_t[i + (i >> 5)];
                offsets[i] = o;
                cursor[i] = o;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = PER * tid + k;
            if (i < n) {
                s_t[i + (i >> 5)] = runc;
                for (int c = runc; c < runc + ch[k] && c < cap_items; c++) item_tile[c] = i;
            }
            runc += ch[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = k * NT + tid;
            if (i < n) chunk_off[i] = s_t[i + (i >> 5)];
        }
    } else {
        // block vb.bx of SCAN_BLOCK tiles (4 consecutive tiles a thread)
        // after the sums of the blocks before it
        const int i0 = vb.bx * SCAN_BLOCK + 4 * tid;
        int ps = 0, pc = 0;
        for (int k = tid; k < vb.bx; k += NT) ps += bsum[2 * k], pc += bsum[2 * k + 1];
        ps = wave_sum(ps), pc = wave_sum(pc);
        if (lane == 0) ws[wv] = ps, wc[wv] = pc;
        __syncthreads();
        if (tid == 0) {
            int a = 0, c = 0;
            for (int k = 0; k < NW; k++) a += ws[k], c += wc[k];
            carry_s = a, carry_c = c;
        }
        __syncthreads();
        int v[4], ch[4];
        for (int k = 0; k < 4; k++) {
            v[k] = i0 + k < n ? counts[i0 + k] : 0;
            ch[k] = item_count(v[k], item_prims);
        }
        const int local = v[0] + v[1] + v[2] + v[3], localc = ch[0] + ch[1] + ch[2] + ch[3];
        const int incl = wave_incl_scan(local, lane), inclc = wave_incl_scan(localc, lane);
        if (lane == 63) ws[wv] = incl, wc[wv] = inclc;
        __syncthreads();
        int before = carry_s, beforec = carry_c;
        for (int k = 0; k < wv; k++) before += ws[k], beforec += wc[k];
        run = before + incl - local, runc = beforec + inclc - localc;
        for (int k = 0; k < 4; k++)
            if (i0 + k < n) {
                offsets[i0 + k] = run;
                cursor[i0 + k] = run;
                chunk_off[i0 + k] = runc;
                for (int c = runc; c < runc + ch[k] && c < cap_items; c++) item_tile[c] = i0 + k;
                counts[i0 + k] = 0;
                run += v[k];
                runc += ch[k];
            }
    }
    if (!whole && vb.bx != vb.gx - 1) return;
    // the last block writes the totals and the frame statistics: unclipped
    // primitives (per-workgroup counts of the setup) plus the clipper's pieces
    if (tid == NT - 1) {
        offsets[n] = run;
        chunk_off[n] = runc;
        cnt->n_pairs = run;
        cnt->n_chunks = runc;
    }
    np = wave_sum(np);
    __syncthreads();  // every wave has read ws[] above
    if (lane == 0) ws[wv] = np;
    __syncthreads();
    if (tid == 0) {
        int a = 0;
        for (int k = 0; k < NW; k++) a += ws[k];
        cnt->n_prims = a;
    }
}
__global__ __launch_bounds__(1024) void k_scan_apply(k_scan_apply_args a) {
    k_scan_apply_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(1024) void k_scan_apply_wide(const k_scan_apply_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_scan_apply_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_scan_apply_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// Scatter each primitive into the lists of the tiles its pixel box touches.
// Slots [0, n_fixed) are one per triangle (per shadow slot); the main view's
// clipped pieces follow them.
struct k_bin_args {
    const uint2* pbox;
    Counters* cnt;
    int32_t tiles_x;
    int32_t shift;  // log2 of the raster tile side
    int32_t tiles_per_view;
    int32_t shadow;
    int32_t* cursor;
    int32_t* list;
    Capacities cap;
    int32_t n_fixed;
    int32_t n_tris;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_bin_body(const k_bin_args& a_, const VB vb) {
    const uint2* __restrict__ pbox = a_.pbox;
    Counters* cnt = a_.cnt;
    const int32_t tiles_x = a_.tiles_x;
    const int32_t shift = a_.shift;
    const int32_t tiles_per_view = a_.tiles_per_view;
    const int32_t shadow = a_.shadow;
    int32_t* __restrict__ cursor = a_.cursor;
    int32_t* __restrict__ list = a_.list;
    const Capacities& cap = a_.cap;
    const int32_t n_fixed = a_.n_fixed;
    const int32_t n_tris = a_.n_tris;

    const int i = vb.bx * blockDim.x + threadIdx.x;
    if (cnt->n_pairs > cap.pairs) {
        if (i == 0) atomicOr(&cnt->overflow, 1);
        return;
    }
    if (cnt->overflow) return;
    // (the clipper's pieces: CLIP_PIECES slots a queued triangle)
    const int np = n_fixed + (shadow ? 0 : min(CLIP_PIECES * min(cnt->n_clipped, cap.clip_tris), cap.clip_subs));
    bool active = false;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0, base = 0;
    if (i < np) {
        const uint2 bb = pbox[i];
        if (bb.y != PRIM_EMPTY) {
            active = true;
            x0 = bb.x & 0xffff, y0 = bb.x >> 16, x1 = bb.y & 0xffff, y1 = bb.y >> 16;
            base = shadow ? (i / n_tris) * tiles_per_view : 0;  // shadow slot s holds [s*Tg, (s+1)*Tg)
        }
    }
    const int tx0 = x0 >> shift, ty0 = y0 >> shift, ntx = active ? (x1 >> shift) - tx0 + 1 : 0;
    const int n = active ? ntx * ((y1 >> shift) - ty0 + 1) : 0;
    if (shadow) {
        // a shadow map gathers a scene in a few of its tiles, so many of the
        // wave's pairs share a tile: one returning atomic per distinct tile
        for (int j = 0; __ballot(j < n) != 0; j++) {
            const bool act = j < n;
            const int key = act ? base + (ty0 + j / ntx) * tiles_x + tx0 + j % ntx : 0;
            const int pos = wave_alloc(cursor, key, act);
            if (act) list[pos] = i;
        }
        return;
    }
    // the main view's pairs are few and spread: the wave's (primitive, tile)
    // pairs go over its lanes, each lane's cursor atomic issued at once (not
    // one returning atomic per distinct tile after another)
    __shared__ int32_t mark[256];
    const int lane = threadIdx.x & 63;
    wave_spread(n, mark + (threadIdx.x & ~63), [&](int o, int j, bool valid) {
        const int otx0 = __shfl(tx0, o), oty0 = __shfl(ty0, o), ontx = __shfl(ntx, o), obase = __shfl(base, o);
        if (!valid) return;
        const int key = obase + (oty0 + j / ontx) * tiles_x + otx0 + j % ontx;
        list[atomicAdd(&cursor[key], 1)] = i - lane + o;
    });
}
__global__ __launch_bounds__(256) void k_bin(k_bin_args a) {
    k_bin_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_bin_wide(const k_bin_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_bin_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_bin_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------------ clipper
// The clipper (clipTriangle, clipping.go:67-157) for the queued triangles:
// Sutherland-Hodgman against the six screen-space planes, a fan of pieces
// with screen-barycentric attributes that keep the parent's recipw
// (raster.go:440-443), appended after the per-triangle primitive slots.
//
// One wave per queued triangle (a frame queues a few thousand at most, so
// the pass is one triangle's latency, not throughput).  Each plane's pass
// runs over the polygon's edges in parallel: lane i takes edge
// (vertex i-1, vertex i) and emits what the serial loop emits for it (the
// intersection and/or vertex i); an exclusive scan of the emit counts
// places them in the serial order, in the wave's other LDS polygon.  Three
// lanes then build each fan piece, and the wave spreads each piece's tile
// counts over its lanes.
constexpr int CLIP_WAVES = 4;   // waves per workgroup (a lane's fan piece needs up to 256 VGPRs)
constexpr int CLIP_VERTS = 12;  // polygon capacity (a triangle clipped by six planes keeps at most 9)

__device__ __forceinline__ int clip_wave(int qi, const ClipJob& jb, const DevScene& s, const ViewConsts& vc,
                                          const float4* __restrict__ tv_scr, const float4* __restrict__ tv_nor,
                                          Prim* __restrict__ prims,
                                          uint2* __restrict__ pbox, ClipRec* __restrict__ clips,
                                          int32_t* __restrict__ tri_clip_slot, int32_t* __restrict__ tile_count,
                                          Counters* cnt, const Capacities& cap, V4 (*poly)[CLIP_VERTS]) {
    const int lane = threadIdx.x & 63;
    const int g = jb.g, i0 = jb.i0, i1 = jb.i1, i2 = jb.i2;
    // Every load first and every store last: a single wave's pass is a chain
    // of memory latencies, and on this part a wait for a load also waits for
    // the stores issued before it.
    const float4 A = tv_scr[i0], B = tv_scr[i1], Cc = tv_scr[i2];
    // the mesh vertices of the transformed vertices i0, i1, i2
    const uint32_t c0 = s.vcol[i0 + jb.dv], c1 = s.vcol[i1 + jb.dv], c2 = s.vcol[i2 + jb.dv];
    const V4 n0 = ld(tv_nor[i0]), n1 = ld(tv_nor[i1]), n2 = ld(tv_nor[i2]);
    float2 u0 = make_float2(0.0f, 0.0f), u1 = u0, u2 = u0;  // UVs exist in scenes with textures only
    if (s.n_textures > 0) u0 = s.vuv[i0 + jb.dv], u1 = s.vuv[i1 + jb.dv], u2 = s.vuv[i2 + jb.dv];
    const V4 p1 = v4(A.x, A.y, A.z, 1), p2 = v4(B.x, B.y, B.z, 1), p3 = v4(Cc.x, Cc.y, Cc.z, 1);
    const float w = vc.vpbox_max_x, h = vc.vpbox_max_y;
    if (lane == 0) poly[0][0] = p1, poly[0][1] = p2, poly[0][2] = p3;
    wave_lds_sync();
    // clipping.go:31-65, the planes in its order
    const Plane planes[6] = {{v4(w, 0, 0, 1), v4(-1, 0, 0, 1)}, {v4(0, 0, 0, 1), v4(1, 0, 0, 1)},
                             {v4(0, h, 0, 1), v4(0, -1, 0, 1)}, {v4(0, 0, 0, 1), v4(0, 1, 0, 1)},
                             {v4(0, 0, 1, 1), v4(0, 0, -1, 1)}, {v4(0, 0, -1, 1), v4(0, 0, 1, 1)}};
    // A plane with all three vertices in front of it by a margin far above
    // the rounding of the intersection points (2^-12 of the largest
    // coordinate, plus 2^-10) keeps every vertex of every polygon the
    // earlier planes make in front of it: those lie on the triangle's edges
    // and chords up to a rounding of a few ulps of the coordinates (the
    // planes are axis-aligned, so each intersection's numerator and
    // denominator carry one or two roundings), so its pass would emit the
    // polygon unchanged, in the same order: skipped.
    float big = 0.0f;
    big = fmaxf(big, fmaxf(fmaxf(__builtin_fabsf(p1.x), __builtin_fabsf(p1.y)), __builtin_fabsf(p1.z)));
    big = fmaxf(big, fmaxf(fmaxf(__builtin_fabsf(p2.x), __builtin_fabsf(p2.y)), __builtin_fabsf(p2.z)));
    big = fmaxf(big, fmaxf(fmaxf(__builtin_fabsf(p3.x), __builtin_fabsf(p3.y)), __builtin_fabsf(p3.z)));
    const float margin = big * 0x1p-12f + 0x1p-10f;
    auto clear_of = [&](const Plane& p) {
        return (dot(sub(p1, p.pos), p.nor) > margin) & (dot(sub(p2, p.pos), p.nor) > margin) &
               (dot(sub(p3, p.pos), p.nor) > margin);
    };
    int na = 3, cur = 0;
#pragma unroll
    for (int pi = 0; pi < 6; pi++) {
        if (na == 0) break;
        if (__builtin_amdgcn_readfirstlane(clear_of(planes[pi]))) continue;
        const V4* a = poly[cur];
        V4* b = poly[cur ^ 1];
        int emit = 0;
        V4 o0 = v4(0, 0, 0, 0), o1 = o0;
        if (lane < na) {
            const V4 e = a[lane], sv = a[lane == 0 ? na - 1 : lane - 1];
            const bool fe = in_front(planes[pi], e), fs = in_front(planes[pi], sv);
            if (fe) {
                if (!fs) o0 = isect(planes[pi], sv, e), o1 = e, emit = 2;
                else o0 = e, emit = 1;
            } else if (fs) {
                o0 = isect(planes[pi], sv, e), emit = 1;
            }
        }
        int incl = emit;
        for (int off = 1; off < 16; off <<= 1) {
            const int u = __shfl_up(incl, off);
            if (lane >= off) incl += u;
        }
        const int pos = incl - emit;
        const int total = __shfl(incl, na - 1);
        if (emit > 0 && pos < CLIP_VERTS) b[pos] = o0;
        if (emit > 1 && pos + 1 < CLIP_VERTS) b[pos + 1] = o1;
        wave_lds_sync();
        na = min(total, CLIP_VERTS);
        cur ^= 1;
    }
    // the job's CLIP_PIECES slots: its pieces, the rest empty
    const int nsub = max(na - 2, 0);
    const int base = qi * CLIP_PIECES;
    if (base + CLIP_PIECES > cap.clip_subs) {
        if (lane == 0) atomicOr(&cnt->overflow, 1);
        return 0;
    }
    // lane 3p + k builds vertex k of fan piece p = (0, p+1, p+2) (one vertex's
    // code a lane); lane p then boxes piece p
    const V4* cl = poly[cur];
    const bool vert = lane < 3 * nsub;
    const int pc = lane / 3, k = lane - 3 * pc;
    V4 q = v4(0, 0, 0, 0), nq = q, m = q;
    float2 uv = make_float2(0.0f, 0.0f);
    uint32_t col = 0;
    if (vert) {
        const V4 cv = cl[k == 0 ? 0 : pc + k];
        const TriSetup par = tri_setup(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
        float bc[3];
        barycoord(par, cv.x, cv.y, bc);
        q = v4(bc[0] * p1.x + bc[1] * p2.x + bc[2] * p3.x, bc[0] * p1.y + bc[1] * p2.y + bc[2] * p3.y,
               bc[0] * p1.z + bc[1] * p2.z + bc[2] * p3.z, 1);
        nq = v4(bc[0] * n0.x + bc[1] * n1.x + bc[2] * n2.x, bc[0] * n0.y + bc[1] * n1.y + bc[2] * n2.y,
                bc[0] * n0.z + bc[1] * n1.z + bc[2] * n2.z, 0);
        uint8_t cc[4];
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++)
            cc[q4] = f2u8(clamp(bc[0] * chan(c0, q4) + bc[1] * chan(c1, q4) + bc[2] * chan(c2, q4), 0, 0xff));
        col = u8x4(cc[0], cc[1], cc[2], cc[3]);
        m = apply(apply(apply(q, vc.vpinv), vc.pinv), vc.vinv);
        uv = make_float2(bc[0] * u0.x + bc[1] * u1.x + bc[2] * u2.x, bc[0] * u0.y + bc[1] * u1.y + bc[2] * u2.y);
    }
    V4 sv[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int src = min(3 * lane + j, 63);
        sv[j] = v4(__shfl(q.x, src), __shfl(q.y, src), __shfl(q.z, src), 1);
    }
    bool ok = false;
    int bx0 = 0, by0 = 0, bx1 = 0, by1 = 0;
    if (lane < nsub)
        ok = pixel_box(sv[0].x, sv[0].y, sv[0].z, sv[1].x, sv[1].y, sv[1].z, sv[2].x, sv[2].y, sv[2].z, vc.W, vc.H,
                       bx0, by0, bx1, by1);
    // each piece's tiles, spread over the wave
    for (int p = 0; p < nsub; p++) {
        if (!__shfl((int)ok, p)) continue;
        const int tx0 = __shfl(bx0, p) >> vc.rt_shift, ty0 = __shfl(by0, p) >> vc.rt_shift;
        const int ntx = (__shfl(bx1, p) >> vc.rt_shift) - tx0 + 1, nty = (__shfl(by1, p) >> vc.rt_shift) - ty0 + 1;
        for (int j = lane; j < ntx * nty; j += 64) {
            const int ty = j / ntx;
            atomicAdd(&tile_count[(ty0 + ty) * vc.rtiles_x + tx0 + (j - ty * ntx)], 1);
        }
    }
    // the stores
    if (lane == 0) tri_clip_slot[g] = base;
    if (vert) {
        // (vertex k keeps the parent's recipw of its own index, as the fan does)
        ClipRec* rp = clips + base + pc;
        rp->pos[k] = make_float4(q.x, q.y, q.z, k == 0 ? A.w : k == 1 ? B.w : Cc.w);
        rp->nor[k] = make_float4(nq.x, nq.y, nq.z, bitsf(col));
        rp->wp[k] = make_float4(m.x, m.y, m.z, 0.0f);
        rp->uv[k] = uv;
    }
    if (lane < CLIP_PIECES) {
        const size_t ps = (size_t)s.n_tris + base + lane;
        if (ok) store_prim(prims, pbox, ps, sv[0], sv[1], sv[2], (uint32_t)g * 8u + (uint32_t)lane, bx0, by0, bx1, by1);
        else store_empty(pbox, ps);
    }
    return nsub;
}

// Waves stride over the queue (its length is known on the device only; the
// grid stays small).  Each workgroup writes its pieces to blk_pieces[its
// index] for the frame statistics (the tile scan sums them): a counter
// that every queued triangle hit would serialise them.
struct k_clip_args {
    DevScene s;
    ViewConsts vc;
    const float4* tv_scr;
    const float4* tv_nor;
    const ClipJob* clip_list;
    Prim* prims;
    uint2* pbox;
    ClipRec* clips;
    int32_t* tri_clip_slot;
    int32_t* tile_count;
    int32_t* blk_pieces;
    Counters* cnt;
    Capacities cap;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_clip_body(const k_clip_args& a_, const VB vb) {
    __shared__ V4 s_poly[CLIP_WAVES][2][CLIP_VERTS];
    __shared__ int32_t s_pieces[CLIP_WAVES];
    Counters* cnt = a_.cnt;
    const int wv = threadIdx.x >> 6;
    // the wave's first job is read before the queue's length is known (a
    // slot within the queue's capacity either way; used only if queued)
    const int q0 = __builtin_amdgcn_readfirstlane(vb.bx * CLIP_WAVES + wv);
    ClipJob jb = a_.clip_list[min(q0, a_.cap.clip_tris - 1)];
    int pieces = 0;
    if (!cnt->overflow) {
        const int n = min(cnt->n_clipped, a_.cap.clip_tris);
        for (int q = q0; q < n; q += vb.gx * CLIP_WAVES) {
            if (q != q0) jb = a_.clip_list[q];
            pieces += clip_wave(q, jb, a_.s, a_.vc, a_.tv_scr, a_.tv_nor, a_.prims, a_.pbox, a_.clips,
                                a_.tri_clip_slot, a_.tile_count, cnt, a_.cap, s_poly[wv]);
        }
    }
    if ((threadIdx.x & 63) == 0) s_pieces[wv] = pieces;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < CLIP_WAVES; k++) t += s_pieces[k];
        a_.blk_pieces[vb.bx] = t;
    }
}
__global__ __launch_bounds__(CLIP_WAVES * 64) void k_clip(k_clip_args a) {
    k_clip_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(CLIP_WAVES * 64) void k_clip_wide(const k_clip_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_clip_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_clip_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------------- rasters
namespace {
// A primitive staged in LDS: the hoisted Barycoord terms, the three depths,
// the order key and its pixel rectangle inside the tile.
struct alignas(16) LPrim {
    float ax, ay, bx, by;
    float abx, aby, acx, acy;
    float bcx, bcy, sabc, az;
    float bz, cz;
    uint32_t key;
    uint32_t rect;  // rx0 | ry0<<5 | (w-1)<<10 | ceil(65536 / w)<<15 (tile-local)
};

// Loads one primitive and intersects its pixel box with the T x T tile;
// returns the candidate count.
template <int T>
__device__ __forceinline__ int stage_prim(const Prim& p, int px0, int py0, LPrim& lp) {
    uint32_t lo = fbits(p.p2.z), hi = fbits(p.p2.w);
    int x0 = max((int)(lo & 0xffff), px0), y0 = max((int)(lo >> 16), py0);
    int x1 = min((int)(hi & 0xffff), px0 + T - 1), y1 = min((int)(hi >> 16), py0 + T - 1);
    if (x0 > x1 || y0 > y1) return 0;
    TriSetup t = tri_setup(p.p0.x, p.p0.y, p.p0.w, p.p1.x, p.p1.z, p.p1.w);
    lp.ax = t.ax, lp.ay = t.ay, lp.bx = t.bx, lp.by = t.by;
    lp.abx = t.abx, lp.aby = t.aby, lp.acx = t.acx, lp.acy = t.acy;
    lp.bcx = t.bcx, lp.bcy = t.bcy, lp.sabc = t.sabc;
    lp.az = p.p0.z, lp.bz = p.p1.y, lp.cz = p.p2.x;
    lp.key = fbits(p.p2.y);
    const int w = x1 - x0 + 1, h = y1 - y0 + 1;
    lp.rect = (uint32_t)(x0 - px0) | ((uint32_t)(y0 - py0) << 5) | ((uint32_t)(w - 1) << 10) |
              (((65536u + (uint32_t)w - 1u) / (uint32_t)w) << 15);
    return w * h;
}

// Candidate idx -> (primitive j, tile-local pixel).  l / w by multiply-shift
// is exact for l < 1024, w <= 32: the product overshoots l / w by less than
// l / 65536 < 1/64, and l / w's fraction is at most 1 - 1/w.
__device__ __forceinline__ void candidate(const int32_t* start, const LPrim* sp, int idx, int& j, int& lx, int& ly) {
    // the last primitive with start <= idx (empty ranges are skipped), in
    // eight fixed steps
    int lo = 0;
#pragma unroll
    for (int step = 128; step >= 1; step >>= 1)
        if (start[lo + step] <= idx) lo += step;
    j = lo;
    int l = idx - start[j];
    uint32_t r = sp[j].rect;
    int w = ((r >> 10) & 31) + 1;
    int q = (int)(mul_u24((uint32_t)l, r >> 15) >> 16);
    lx = (int)(r & 31) + (l - q * w);
    ly = (int)((r >> 5) & 31) + q;
}

__device__ __forceinline__ bool eval_candidate(const LPrim& p, float px, float py, float& z) {
    TriSetup t;
    t.ax = p.ax, t.ay = p.ay, t.bx = p.bx, t.by = p.by;
    t.abx = p.abx, t.aby = p.aby, t.acx = p.acx, t.acy = p.acy;
    t.bcx = p.bcx, t.bcy = p.bcy, t.sabc = p.sabc;
    return cover_z(t, px, py, p.az, p.bz, p.cz, z);  // raster.go:487-496
}
}  // namespace

// Work items of the rasters: (tile, up to ITEM_PRIMS primitives of its
// list), so that a tile a scene crowds -- a shadow map fitted to the whole
// camera frustum concentrates the scene in a few hundred of its tiles -- is
// spread over many workgroups.  k_scan_apply lists each item's tile.
// Persistent workgroups take items i = blockIdx.x, blockIdx.x + gridDim.x, ...
// Each item evaluates its candidates (pixel box x tile) exactly as
// drawClipped / drawDepth do, reduces them in LDS and merges the tile's
// covered pixels into the target with global atomics (order-independent,
// and with what the setup drew and the tile's other items):
//   MODE_VIS    64-bit min of (~ordered(z) << 32 | triangle*8+sub): largest
//               depth wins, ties go to the first primitive in sequential
//               order -- FragmentBuffer.DepthTest/Set under Workers(1)
//               (buffer/buffer.go:221-280); a NaN depth is flagged;
//   MODE_EXACT  also the first fragment and the first NaN fragment of each
//               pixel, for the NaN rule (a NaN depth only survives as the
//               very first fragment of a pixel);
//   MODE_SHADOW the running maximum of render/shadow.go:214-228 (maps zeroed
//               once, never cleared) as a max over ordered depth bits.
enum { MODE_VIS = 0, MODE_EXACT = 1, MODE_SHADOW = 2 };

struct k_raster_args {
    const Prim* prims;
    const int32_t* tile_off;
    const int32_t* chunk_off;
    const int32_t* item_tile;
    const int32_t* tile_list;
    Counters* cnt;
    RasterTarget rt;
    int32_t cap_items;
    int32_t item_prims;
    int32_t gx, gy;  // the grid of this launch
};
template <int MODE, int T>
__device__ __forceinline__ void k_raster_body(const k_raster_args& a_, const VB vb) {
    const int32_t item_prims = a_.item_prims;
    const Prim* __restrict__ prims = a_.prims;
    const int32_t* __restrict__ tile_off = a_.tile_off;
    const int32_t* __restrict__ chunk_off = a_.chunk_off;
    const int32_t* __restrict__ item_tile = a_.item_tile;
    const int32_t* __restrict__ tile_list = a_.tile_list;
    Counters* cnt = a_.cnt;
    const RasterTarget& rt = a_.rt;
    const int32_t cap_items = a_.cap_items;

    constexpr bool SHADOW = MODE == MODE_SHADOW, EXACT = MODE == MODE_EXACT;
    constexpr int TP = T * T;  // pixels of a tile: 64, 256 or 1024 (a quarter, one or four per thread)
    __shared__ LPrim sp[CHUNK];
    __shared__ int32_t start[CHUNK + 1];
    __shared__ int32_t wsum[4];
    __shared__ unsigned long long svis[SHADOW ? 1 : TP];
    __shared__ uint32_t smax[SHADOW ? TP : 1];
    __shared__ uint32_t sfirst[EXACT ? TP : 1];
    __shared__ uint32_t snan[EXACT ? TP : 1];
    const int tid = threadIdx.x;
    const int n_items = cnt->n_chunks;
    if (cnt->overflow || n_items > cap_items) return;
    bool saw_nan = false;
    for (int item = vb.bx; item < n_items; item += vb.gx) {
        const int tile = item_tile[item];
        const int slot = SHADOW ? tile / rt.tiles_per_map : 0;
        const int t = SHADOW ? tile % rt.tiles_per_map : tile;
        const int px0 = (t % rt.tiles_x) * T, py0 = (t / rt.tiles_x) * T;
        const int first_item = chunk_off[tile];
        const int beg = tile_off[tile] + (item - first_item) * item_prims;
        const int n = min(item_prims, tile_off[tile + 1] - beg);
        for (int pp = tid; pp < TP; pp += 256) {
            if (SHADOW) smax[pp] = depth_ord(0.0f);
            else svis[pp] = VIS_EMPTY;
            if (EXACT) sfirst[pp] = 0xffffffffu, snan[pp] = 0xffffffffu;
        }
        for (int b0 = 0; b0 < n; b0 += CHUNK) {
            int c = 0;
            if (b0 + tid < n) c = stage_prim<T>(prims[tile_list[beg + b0 + tid]], px0, py0, sp[tid]);
            block_scan_256(c, start, wsum);
            const int total = start[CHUNK];
            for (int idx = tid; idx < total; idx += 256) {
                int j, lx, ly;
                candidate(start, sp, idx, j, lx, ly);
                float z;
                if (!eval_candidate(sp[j], (float)(px0 + lx) + 0.5f, (float)(py0 + ly) + 0.5f, z)) continue;
                const int pix = ly * T + lx;
                const uint32_t key = sp[j].key;
                if (EXACT) atomicMin(&sfirst[pix], key);
                if (z != z) {
                    saw_nan = true;
                    if (EXACT) atomicMin(&snan[pix], key);
                    continue;
                }
                if (SHADOW) atomicMax(&smax[pix], depth_ord(z));
                else atomicMin(&svis[pix], ((unsigned long long)depth_key_hi(z) << 32) | key);
            }
            __syncthreads();
        }
        for (int pp = tid; pp < TP; pp += 256) {
            const int x = px0 + (pp & (T - 1)), y = py0 + pp / T;
            // (the setup may have drawn small triangles into the tile, and
            // other items of it merge here too: an atomic, which needs no
            // read of the old value first)
            if (SHADOW) {
                if (x < rt.W && y < rt.H && smax[pp] > depth_ord(0.0f))
                    atomicMax(&rt.smap_ord[(size_t)slot * rt.W * rt.H + (size_t)y * rt.W + x], smax[pp]);
            } else if (x < rt.W && y < rt.H) {
                const size_t o = (size_t)y * rt.W + x;
                if (svis[pp] != VIS_EMPTY) atomicMin(&rt.vis[o], svis[pp]);
                if (EXACT) {
                    if (sfirst[pp] != 0xffffffffu) atomicMin(&rt.first_all[o], sfirst[pp]);
                    if (snan[pp] != 0xffffffffu) atomicMin(&rt.first_nan[o], snan[pp]);
                }
            }
        }
        __syncthreads();
    }
    if (saw_nan) atomicOr(SHADOW ? &cnt->shadow_nan : &cnt->nan_flag, 1);
}
template <int MODE, int T>
__global__ __launch_bounds__(256) void k_raster(k_raster_args a) {
    k_raster_body<MODE, T>(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
template <int MODE, int T>
__global__ __launch_bounds__(256) void k_raster_wide(const k_raster_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_raster_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_raster_body<MODE, T>(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// Fragment depth per pixel from the visibility key (-1 where uncovered, as
// ao.go reads an empty fragment); after the exact-order pass, a NaN-depth
// fragment that came first at its pixel takes it.
struct k_resolve_args {
    RasterTarget rt;
    Counters* cnt;
    float* depth;
    DevScene s;
    int32_t* ao_xs;   // scenes with AO: each row's AO pixels, left to right (for k_ao)
    int32_t* ao_cnt;  // and their count
    const int4* tri_rec;
    int32_t gx, gy;  // the grid of this launch
};
// One workgroup per row, 2048 pixels at a time (eight of 256 a thread, all
// loads first): the depth, and with AO the row's list of the pixels k_ao
// computes -- those whose material has AmbientOcclusion, and pixel (0,0)
// when covered (uncovered pixels are shaded from it) -- placed by wave
// ballots and a scan of the 32 wave counts.
constexpr int RESOLVE_PER = 8;
template <bool EXACT>
__device__ __forceinline__ void k_resolve_body(const k_resolve_args& a_, const VB vb) {
    const RasterTarget& rt = a_.rt;
    Counters* cnt = a_.cnt;
    float* __restrict__ depth = a_.depth;
    const DevScene& s = a_.s;
    int32_t* __restrict__ ao_xs = a_.ao_xs;
    const int4* __restrict__ tri_rec = a_.tri_rec;
    __shared__ int32_t s_cnt[RESOLVE_PER * 4 + 1];

    const int y = vb.bx, W = rt.W;
    if (y >= rt.H || cnt->overflow) return;
    const int tid = threadIdx.x, wv = tid >> 6;
    const size_t row = (size_t)y * W;
    int running = 0;  // the row's AO pixels so far
    for (int x0 = 0; x0 < W; x0 += 256 * RESOLVE_PER) {
        unsigned long long v[RESOLVE_PER];
#pragma unroll
        for (int k = 0; k < RESOLVE_PER; k++) {
            const int x = x0 + 256 * k + tid;
            v[k] = x < W ? rt.vis[row + x] : VIS_EMPTY;
        }
        bool need[RESOLVE_PER];
#pragma unroll
        for (int k = 0; k < RESOLVE_PER; k++) {
            const int x = x0 + 256 * k + tid;
            need[k] = false;
            if (x >= W) continue;
            const size_t o = row + x;
            if (EXACT) {
                const uint32_t fn = rt.first_nan[o];
                if (fn != 0xffffffffu && fn == rt.first_all[o]) {
                    v[k] = ((unsigned long long)KEY_HI_NAN << 32) | fn;
                    rt.vis[o] = v[k];
                }
            }
            const uint32_t hi = (uint32_t)(v[k] >> 32);
            const float d = v[k] == VIS_EMPTY ? -1.0f : (hi == KEY_HI_NAN ? __builtin_nanf("") : depth_from_key_hi(hi));
            depth[o] = d;
            if (!(__builtin_fabsf(d) <= 3.40282347e38f)) atomicOr(&cnt->depth_nonfinite, 1);  // lets k_ao skip NaN checks
            if (ao_xs && v[k] != VIS_EMPTY) {
                if (o == 0) {
                    need[k] = true;
                } else {
                    const int mat = tri_rec[(uint32_t)v[k] >> 3].w;
                    need[k] = mat >= 0 && s.mats[mat].ambient_occlusion;
                }
            }
        }
        if (!ao_xs) continue;
        uint64_t bal[RESOLVE_PER];
#pragma unroll
        for (int k = 0; k < RESOLVE_PER; k++) {
            bal[k] = __ballot(need[k]);
            if ((tid & 63) == 0) s_cnt[4 * k + wv] = __popcll(bal[k]);
        }
        __syncthreads();
        if (tid == 0) {  // exclusive scan, pass-major then wave: the pixels' order
            int acc = 0;
            for (int j = 0; j < RESOLVE_PER * 4; j++) {
                const int c = s_cnt[j];
                s_cnt[j] = acc;
                acc += c;
            }
            s_cnt[RESOLVE_PER * 4] = acc;
        }
        __syncthreads();
        const uint64_t below = (1ull << (tid & 63)) - 1ull;
#pragma unroll
        for (int k = 0; k < RESOLVE_PER; k++)
            if (need[k]) ao_xs[row + running + s_cnt[4 * k + wv] + __popcll(bal[k] & below)] = x0 + 256 * k + tid;
        running += s_cnt[RESOLVE_PER * 4];
        __syncthreads();  // (s_cnt is rewritten by the next 2048 pixels)
    }
    if (ao_xs && tid == 0) a_.ao_cnt[y] = running;
}
template <bool EXACT>
__global__ __launch_bounds__(256) void k_resolve(k_resolve_args a) {
    k_resolve_body<EXACT>(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
template <bool EXACT>
__global__ __launch_bounds__(256) void k_resolve_wide(const k_resolve_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_resolve_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_resolve_body<EXACT>(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ---------------------------------------------------------------- shade
namespace {

struct GFrag {
    float depth;
    V4 nor, facenor, wpos;
    uint32_t col;
    int mat;  // flat material id (-1 when matAt gives nil)
    float u, v, du, dv;  // texture coordinates and their screen derivatives (scenes with textures)
};

// Rebuilds the G-buffer fragment of pixel (x, y) from its visibility key,
// as drawClipped wrote it (render/raster.go:487-568): Barycoord at the pixel
// centre, linear depth, perspective-correct vertex colour, unit
// interpolated normal, face normal and interpWorldPos of the back-transformed
// vertices; in a scene with textures also the perspective-correct UV and its
// squared screen derivatives Du, Dv from Barycoords one pixel right and one
// down (render/raster.go:511-535; their normalisation is unconditional).
__device__ GFrag fetch_fragment(const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b, int x, int y,
                                uint64_t key) {
    const uint32_t ord = (uint32_t)key;
    const int g = (int)(ord >> 3), piece = (int)(ord & 7);
    const int4 rec = b.tri_rec[g];  // the setup's record: vertices and material
    V4 P[3], N[3], M[3];
    float rw[3];
    uint32_t C3[3];
    float2 UV[3];
    const bool tex = s.n_textures > 0;
    if (piece == (int)SUB_DIRECT) {
        const int vid[3] = {rec.x, rec.y, rec.z};
        for (int k = 0; k < 3; k++) {
            const float4 a = b.tv_scr[vid[k]];
            P[k] = v4(a.x, a.y, a.z, 1);
            rw[k] = a.w;
            N[k] = ld(b.tv_nor[vid[k]]);
            M[k] = ld(b.tv_wp[vid[k]]);
            C3[k] = b.tv_col[vid[k]];
            UV[k] = tex ? b.tv_uv[vid[k]] : make_float2(0.0f, 0.0f);
        }
    } else {
        const ClipRec& r = b.clips[b.tri_clip_slot[g] + piece];
        for (int k = 0; k < 3; k++) {
            P[k] = v4(r.pos[k].x, r.pos[k].y, r.pos[k].z, 1);
            rw[k] = r.pos[k].w;
            N[k] = v4(r.nor[k].x, r.nor[k].y, r.nor[k].z, 0);
            C3[k] = fbits(r.nor[k].w);
            M[k] = ld(r.wp[k]);
            UV[k] = r.uv[k];
        }
    }
    TriSetup ts = tri_setup(P[0].x, P[0].y, P[1].x, P[1].y, P[2].x, P[2].y);
    float bc[3];
    barycoord(ts, (float)x + 0.5f, (float)y + 0.5f, bc);
    GFrag f;
    f.depth = bc[0] * P[0].z + bc[1] * P[1].z + bc[2] * P[2].z;
    float wc1 = rw[0] * bc[0], wc2 = rw[1] * bc[1], wc3 = rw[2] * bc[2];
    float norm = 1.0f;
    if (c.v.perspect) norm = 1 / (wc1 + wc2 + wc3);
    uint8_t cc[4];
    for (int k = 0; k < 4; k++)
        cc[k] = f2u8(clamp((wc1 * chan(C3[0], k) + wc2 * chan(C3[1], k) + wc3 * chan(C3[2], k)) * norm, 0, 0xff));
    f.col = u8x4(cc[0], cc[1], cc[2], cc[3]);
    f.nor = unit(v4(bc[0] * N[0].x + bc[1] * N[1].x + bc[2] * N[2].x, bc[0] * N[0].y + bc[1] * N[1].y + bc[2] * N[2].y,
                    bc[0] * N[0].z + bc[1] * N[1].z + bc[2] * N[2].z, 0));
    V4 m1 = v4(M[0].x, M[0].y, M[0].z, 0), m2 = v4(M[1].x, M[1].y, M[1].z, 0), m3 = v4(M[2].x, M[2].y, M[2].z, 0);
    f.facenor = unit(cross(sub(m2, m1), sub(m3, m1)));
    f.wpos = v4(bc[0] * m1.x + bc[1] * m2.x + bc[2] * m3.x, bc[0] * m1.y + bc[1] * m2.y + bc[2] * m3.y,
                bc[0] * m1.z + bc[1] * m2.z + bc[2] * m3.z, 1);
    const int mat = rec.w;
    f.mat = mat;
    f.u = f.v = f.du = f.dv = 0.0f;
    if (tex) {
        f.u = (wc1 * UV[0].x + wc2 * UV[1].x + wc3 * UV[2].x) * norm;
        f.v = (wc1 * UV[0].y + wc2 * UV[1].y + wc3 * UV[2].y) * norm;
        if (mat >= 0) {
            float bx[3], by[3];
            const float px = (float)x + 0.5f, py = (float)y + 0.5f;
            barycoord(ts, px + 1, py, bx);
            const float w1x = rw[0] * bx[0], w2x = rw[1] * bx[1], w3x = rw[2] * bx[2];
            const float nx = 1 / (w1x + w2x + w3x);
            barycoord(ts, px, py + 1, by);
            const float w1y = rw[0] * by[0], w2y = rw[1] * by[1], w3y = rw[2] * by[2];
            const float ny = 1 / (w1y + w2y + w3y);
            const float uU = (w1x * UV[0].x + w2x * UV[1].x + w3x * UV[2].x) * nx;
            const float uX = (w1x * UV[0].y + w2x * UV[1].y + w3x * UV[2].y) * nx;
            const float vV = (w1y * UV[0].x + w2y * UV[1].x + w3y * UV[2].x) * ny;
            const float vY = (w1y * UV[0].y + w2y * UV[1].y + w3y * UV[2].y) * ny;
            f.du = (uU - f.u) * (uU - f.u) + (uX - f.v) * (uX - f.v);
            f.dv = (vV - f.u) * (vV - f.u) + (vY - f.v) * (vY - f.v);
        }
    }
    return f;
}

// ---- textures (buffer/texture.go:83-195)
// image.RGBA.RGBAAt of one level: transparent black outside it (Go ints:
// int(Floor(NaN)) is the most negative one, outside)
__device__ __forceinline__ uint32_t texel(const DevScene& s, const DevTexture& t, int lv, int64_t x, int64_t y) {
    if (x < 0 || y < 0 || x >= t.w[lv] || y >= t.h[lv]) return 0u;
    return s.tex_data[t.off[lv] + y * t.w[lv] + x];
}
// math.LerpC (math/interpolate.go:55-62)
__device__ __forceinline__ uint32_t lerp_c(uint32_t a, uint32_t b, float t) {
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) o |= (uint32_t)f2u8(chan(a, k) + t * (chan(b, k) - chan(a, k))) << (8 * k);
    return o;
}
// queryBilinear, buffer/texture.go:162-195
__device__ uint32_t tex_bilinear(const DevScene& s, const DevTexture& t, int lv, float u, float v) {
    const int dx = t.w[lv], dy = t.h[lv];
    if (dx == 1 && dy == 1) return texel(s, t, lv, 0, 0);
    const float x = u * ((float)dx - 1), y = v * ((float)dy - 1);
    const float x0 = floorf(x), y0 = floorf(y);
    const int64_t i = f2i(x0), j = f2i(y0);
    const uint32_t p1 = texel(s, t, lv, i, j);
    const uint32_t p2 = i < dx - 1 ? texel(s, t, lv, i + 1, j) : p1;
    const uint32_t q1 = lerp_c(p1, p2, x - x0);
    const uint32_t p3 = j < dy - 1 ? texel(s, t, lv, i, j + 1) : p1;
    const uint32_t p4 = (i < dx - 1 && j < dy - 1) ? texel(s, t, lv, i + 1, j + 1) : p1;
    const uint32_t q2 = lerp_c(p3, p4, x - x0);
    return lerp_c(q1, q2, y - y0);
}
// Texture.Query, buffer/texture.go:83-154
__device__ uint32_t tex_query(const DevScene& s, const DevTexture& t, float lod, float u, float v) {
    double iu, iv;  // math.Modf: Modf(+-Inf) = +-Inf, NaN
    float fu = isinf(u) ? __builtin_nanf("") : (float)modf((double)u, &iu);
    float fv = isinf(v) ? __builtin_nanf("") : (float)modf((double)v, &iv);
    if (isinf(u)) iu = u;
    if (isinf(v)) iv = v;
    if ((float)iu != 0 && fu == 0) fu = 1;
    if (fu < 0) fu = 1 - fu;
    if ((float)iv != 0 && fv == 0) fv = 1;
    if (fv < 0) fv = 1 - fv;
    if (!t.use_mipmap) {  // queryL0, texture.go:130-141
        if (t.w[0] == 1 && t.h[0] == 1) return texel(s, t, 0, 0, 0);
        return texel(s, t, 0, f2i(floorf(fu * ((float)t.w[0] - 1))), f2i(floorf(fv * ((float)t.h[0] - 1))));
    }
    if (lod < 0) lod = 0;
    else if (lod >= (float)t.n) lod = (float)(t.n - 1);
    if (lod <= 1) return tex_bilinear(s, t, 0, fu, fv);
    lod -= 1;
    const int h = (int)f2i(floorf(lod)), l = h + 1;  // lod is within [1, n-1] here
    if (l >= t.n) return tex_bilinear(s, t, h, fu, fv);
    const float p = lod - (float)h;
    if (approx_eq(p, 0, 1e-7f)) return tex_bilinear(s, t, h, fu, fv);
    return lerp_c(tex_bilinear(s, t, h, fu, fv), tex_bilinear(s, t, l, fu, fv), p);
}
// math.Log2 as Go computes it: Frexp, then Log of the fraction
__device__ __forceinline__ double go_log2(double x) {
    if (!(x < INFINITY)) return x;  // Log2(+Inf) = +Inf, Log2(NaN) = NaN
    int e;
    const double f = frexp(x, &e);
    if (f == 0.5) return (double)(e - 1);
    return log(f) * (1 / 0.693147180559945309417232121458176568) + (double)e;
}
// FragmentShader's texture colour (shader/blinn_cpu.go:28-35)
__device__ uint32_t tex_fragment(const DevScene& s, const DevTexture& t, const GFrag& f) {
    float lod = 0.0f;
    if (t.use_mipmap) {
        float siz = (float)t.size * sqrt32(max2(f.du, f.dv));
        if (siz < 1) siz = 1;
        lod = (float)go_log2((double)siz);
    }
    return tex_query(s, t, lod, f.u, 1 - f.v);
}

// shader.FragmentShader (shader/blinn_cpu.go:24-106).  Returns false on the
// no-lights early return (texture colour, no linear sums).
__device__ bool fragment_shader(const DevScene& s, const ShadeConsts& c, const DevMaterial& m, const GFrag& f,
                                uint32_t& out, float lin[3]) {
    // the texture colour: the uniform one, or Query at the fragment's UV
    const uint32_t tc = m.tex_id >= 0 ? tex_fragment(s, s.textures[m.tex_id], f) : m.tex_rgba;
    if (s.n_lights == 0) {
        out = tc;
        return false;
    }
    float tex[3], la[3];
    if (m.tex_id >= 0) {  // La = sum_e Intensity_e * col (blinn_cpu.go:44-52)
        for (int k = 0; k < 3; k++) {
            tex[k] = chan(tc, k);
            la[k] = 0.0f;
            for (int e = 0; e < s.n_ambient; e++) la[k] += s.ambient[e] * tex[k];
        }
    } else {
        for (int k = 0; k < 3; k++) tex[k] = m.tex[k], la[k] = m.la[k];
    }
    float LdR = 0, LdG = 0, LdB = 0, LsR = 0, LsG = 0, LsB = 0;
    const V4 n = m.flat_shading ? f.facenor : f.nor;
    const V4 x = f.wpos;
    const V4 V = unit(sub(v4(c.v.cam[0], c.v.cam[1], c.v.cam[2], 1), x));
    for (int i = 0; i < s.n_lights; i++) {
        const DevLight& l = s.lights[i];
        V4 L;
        float I;
        if (l.kind == 0) {
            V4 ldir = sub(v4(l.pos[0], l.pos[1], l.pos[2], 1), x);
            L = unit(ldir);
            I = l.intensity / len(ldir);
        } else {
            L = scale(v4(l.dir[0], l.dir[1], l.dir[2], 0), -1, -1, -1, 1);
            I = l.intensity;
        }
        V4 H = unit(add(L, V));
        float Ld = clamp(dot(n, L), 0, 1);
        float Ls = pow32_fast(clamp(dot(n, H), 0, 1), m.shininess);  // = (float)pow64(double, double)
        LdR += Ld * tex[0] * I;
        LdG += Ld * tex[1] * I;
        LdB += Ld * tex[2] * I;
        LsR += Ls * l.color[0] * I;
        LsG += Ls * l.color[1] * I;
        LsB += Ls * l.color[2] * I;
    }
    float r = la[0] + (m.diffuse[0] * LdR / 255.0f) + (m.specular[0] * LsR / 255.0f);
    float g = la[1] + (m.diffuse[1] * LdG / 255.0f) + (m.specular[1] * LsG / 255.0f);
    float bb = la[2] + (m.diffuse[2] * LdB / 255.0f) + (m.specular[2] * LsB / 255.0f);
    lin[0] = r, lin[1] = g, lin[2] = bb;
    out = u8x4(f2u8(clamp(round32(r), 0, 0xff)), f2u8(clamp(round32(g), 0, 0xff)), f2u8(clamp(round32(bb), 0, 0xff)),
               (uint8_t)(tc >> 24));
    return true;
}

// render/shadow.go:230-283 summed over all lights (raster.go:339-347): the
// number of lights for which shadingVisibility reports true.
// apply() of a matrix whose last row is (0 0 0 1), for a finite v: the zero
// terms of the last row's fma64 chain add zeros, so its result is v.w (up to
// the sign of a zero).  With zeros off the diagonal of the upper 3x3 as well
// (orthographic projections, viewports) each row is one fma64.  Nothing the
// shadow test computes from these vectors tells the sign of a zero: pos()
// tests w == 0 before dividing, and int() and the depth comparison treat -0
// as 0.  The coefficients come from a ShadowXf: fma64(a, v, c) with a already
// float64 ((double)a is exact, so the result is the same) and the vector's
// components converted once.
__device__ __forceinline__ bool finite4(V4 v) { return fin(v.x) && fin(v.y) && fin(v.z) && fin(v.w); }
__device__ __forceinline__ float fma64d(double a, double v, float c) { return (float)__builtin_fma(a, v, (double)c); }
__device__ __forceinline__ V4 apply_view_xf(V4 v, const ShadowXf& m) {
    const double x = v.x, y = v.y, z = v.z;
    const double* a = m.view;
    return v4(fma64d(a[0], x, fma64d(a[1], y, fma64d(a[2], z, m.view_c[0] * v.w))),
              fma64d(a[3], x, fma64d(a[4], y, fma64d(a[5], z, m.view_c[1] * v.w))),
              fma64d(a[6], x, fma64d(a[7], y, fma64d(a[8], z, m.view_c[2] * v.w))), v.w);
}
__device__ __forceinline__ V4 apply_diag_xf(V4 v, const double* d, const float* c) {
    return v4(fma64d(d[0], v.x, c[0] * v.w), fma64d(d[1], v.y, c[1] * v.w), fma64d(d[2], v.z, c[2] * v.w), v.w);
}


__device__ __forceinline__ int occluded(const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b, int fx, int fy,
                                        float depth) {
    int occ = 0;
    const int64_t len = (int64_t)c.SW * c.SH;
    // SW and SH exact in float32, y * SW within 32 bits
    const bool small_map = c.SW > 0 && c.SH > 0 && c.SW < (1 << 24) && c.SH < (1 << 24) && len < ((int64_t)1 << 31);
    const float fSW = (float)c.SW, fSH = (float)c.SH;
    // the screen-to-world step is the same for every light
    const V4 qw = apply(v4((float)fx, (float)fy, depth, 1), c.v.s2w);
    // pos()'s reciprocal of w, the same for every light whose transforms
    // keep w (the sparse shape with finite values)
    const bool qw_unit = qw.w == 1 || qw.w == 0;
    const float qw_inv = 1.0f / qw.w;
    for (int i = 0; i < s.n_lights; i++) {
        const DevLight& l = s.lights[i];
        if (!l.cast_shadow) {
            occ++;
            continue;
        }
        V4 q;
        if (c.shadow_sparse && finite4(qw)) {
            const ShadowXf& xf = b.light_xf[i];
            q = apply_view_xf(qw, xf);
            if (finite4(q)) q = apply_diag_xf(q, xf.proj, xf.proj_c);
            else q = apply(q, b.light_proj[i]);
            if (finite4(q)) {
                // w is still qw's: pos() with the pixel's one reciprocal
                q = apply_diag_xf(q, xf.svp, xf.svp_c);
                q = qw_unit ? v4(q.x, q.y, q.z, 1) : v4(q.x * qw_inv, q.y * qw_inv, q.z * qw_inv, 1);
            } else {
                q = pos(apply(q, c.svp));
            }
        } else {
            q = apply(qw, b.light_view[i]);
            q = apply(q, b.light_proj[i]);
            q = pos(apply(q, c.svp));
        }
        // Go's int(x) and int arithmetic (64-bit, wrapping).  Where every
        // lane's texel is inside the map in both coordinates (-1 < x < SW,
        // -1 < y < SH: truncation lands in [0, SW) and [0, SH)), the common
        // case, the wave converts in 32 bits and takes one 24-bit multiply
        // (a uniform branch: the 64-bit conversions and the quarter-rate
        // multiplies of the general path are skipped, not just masked)
        const bool in_map = small_map & (q.x > -1.0f) & (q.x < fSW) & (q.y > -1.0f) & (q.y < fSH);
        int64_t idx;
        if (__ballot(!in_map) == 0) {
            idx = (int64_t)(mul_u24((uint32_t)(int32_t)q.y, (uint32_t)c.SW) + (uint32_t)(int32_t)q.x);
        } else {
            const int64_t lx = f2i(q.x), ly = f2i(q.y);
            idx = (int64_t)((uint64_t)lx + (uint64_t)ly * (uint64_t)(int64_t)c.SW);
        }
        if (idx > 0 && idx < len) {
            float sz = depth_from_ord(b.smaps[(size_t)l.slot * len + idx]);
            if (q.z < sz - 0.03f) occ++;
        }
    }
    return occ;
}

// Correctly rounded float32 sqrt for finite x >= 2^-96: the hardware root
// (within 1 ulp) corrected by the sign of the residuals of its neighbours,
// the expansion the compiler emits for sqrtf minus its small-input scaling
// and special-class handling, which these inputs never need.
__device__ __forceinline__ float sqrt32_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
    const float rdn = __builtin_fmaf(-dn, s, x), rup = __builtin_fmaf(-up, s, x);
    float r = rdn <= 0.0f ? dn : s;
    r = rup > 0.0f ? up : r;
    return r;
}

// AmbientOcclusionShade's last step (material/ao.go:34-37): the fragment
// colour scaled by the AO factor k_ao computed for its pixel.
__device__ __forceinline__ uint32_t ao_apply(uint32_t col, float total) {
    return u8x4(f2u8(total * chan(col, 0)), f2u8(total * chan(col, 1)), f2u8(total * chan(col, 2)), (uint8_t)(col >> 24));
}

// render/raster.go:324-359 for a fragment f (the stored fragment of pixel
// (fx, fy)) shaded with material `mat`; `col` is the fragment's stored colour.
__device__ uint32_t shade_fragment(const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b, const GFrag& f,
                                   int fx, int fy, int mat, uint32_t col, float lin[3], bool& has_lin) {
    has_lin = false;
    if (mat < 0) return col;  // vertex colour; AO with a nil material returns it
    const DevMaterial& m = s.mats[mat];
    has_lin = fragment_shader(s, c, m, f, col, lin);
    if (c.shadow_map && m.receive_shadow) {
        int occ = occluded(s, c, b, fx, fy, f.depth);
        float w = __builtin_ldexpf(1.0f, -occ);  // math.Pow(0.5, visibles), exact
        col = u8x4(f2u8(chan(col, 0) * w), f2u8(chan(col, 1) * w), f2u8(chan(col, 2) * w), (uint8_t)(col >> 24));
    }
    if (m.ambient_occlusion && !b.col_pre) col = ao_apply(col, b.ao[(size_t)fy * c.v.W + fx]);
    return col;
}

__device__ __forceinline__ uint32_t gamma(const ShadeConsts& c, uint32_t col) {
    if (!c.gamma_correct) return col;
    return (uint32_t)c.gamma_lut[col & 0xff] | ((uint32_t)c.gamma_lut[(col >> 8) & 0xff] << 8) |
           ((uint32_t)c.gamma_lut[(col >> 16) & 0xff] << 16) | (col & 0xff000000u);
}

}  // namespace

// Screen-space AO (material/ao.go:20-73) for every pixel whose material has
// it, plus pixel (0,0) (the uncovered pixels are shaded from it).  One
// workgroup = 64 consecutive AO pixels of one row (k_resolve's list) x the 8
// march directions, four waves of two directions each; the lanes hold the
// row's AO pixels left to right, so each tap of every direction reads nearly
// contiguous depths across the wave, and no lane marches for a pixel without
// AO.  Workgroup i takes chunk i % nbx of row i / nbx, so the dispatcher
// (workgroup i runs on XCD i % 8) deals every XCD a share of every part of
// the frame: each XCD gets the same mix of full and partial rows, and of the
// rows near the borders, whose marches take the per-lane path.
// maxElevationAngle returns at the first tap outside the buffer; the tap
// coordinates move monotonically along the ray, so a tap never re-enters once
// out and masking every later tap gives the same result -- which turns the
// march into a fixed 99-step loop with no divergence.  The running Max of atan
// over the taps equals atan of the largest elevation/distance ratio (atan and
// its rounding to float32 are monotone); a NaN ratio poisons it as math.Max
// does.  The per-pixel factor pow(total, 10000) (ao.go:31-33) is written for
// k_shade.
struct k_ao_args {
    ShadeConsts c;
    const int32_t* ao_xs;    // k_resolve's lists
    const int32_t* ao_cnt;   // and counts
    const float* depth;
    float* ao;
    const float* ao_tab;
    int32_t ao_fast;
    const Counters* cnt;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_ao_body(const k_ao_args& a_, const VB vb) {
    const ShadeConsts& c = a_.c;
    const int32_t* __restrict__ ao_xs = a_.ao_xs;
    const int32_t* __restrict__ ao_cnt = a_.ao_cnt;
    const float* __restrict__ depth = a_.depth;
    float* __restrict__ ao = a_.ao;
    const float* __restrict__ ao_tab = a_.ao_tab;
    const int32_t ao_fast = a_.ao_fast;
    const Counters* cnt = a_.cnt;

    __shared__ float s_m[8][64];
    __shared__ float s_win[2][2][AO_WIN];  // the two x-axis marches: each its two rows around the pixels
    __shared__ int32_t s_dbase[4][100];  // diagonal marches, tap t: byte offset of lane 0's tap
    __shared__ double s_drinv[4][100];   // and fl64(1 / distance), where every lane's tap shares them
    // and for every diagonal tap, the row's part (the chunk shares its row):
    // the row's first pixel (negative outside the buffer) and the distance
    // table's row term
    __shared__ int32_t s_drow[4][100], s_dtab[4][100];
    if (cnt->overflow) return;
    const int W = c.v.W, H = c.v.H;
    const int nbx = (W + 63) / 64, nb = nbx * H;
    const int blk = vb.bx;
    if (blk >= nb) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // chunk blk % nbx of row y: the row's AO pixels r0 .. r0+63; past the
    // row's count the workgroup ends here, all its waves alike
    const int y = blk / nbx, r0 = (blk % nbx) * 64;
    const int n_row = ao_cnt[y];
    if (r0 >= n_row) return;
    const int32_t* xs = ao_xs + (size_t)y * W;
    const bool need = r0 + lane < n_row;  // lanes past the row's last AO pixel compute nothing anyone reads
    const int x = xs[min(r0 + lane, n_row - 1)];
    const int x_first = xs[r0], x_last = xs[min(r0 + 63, n_row - 1)];  // (uniform)
    const int li = x - x_first;  // the lane's column counted from the chunk's first pixel
    const int span = x_last - x_first;
    // the x marches read a window of the row staged in LDS when it spans
    // every tap of the chunk (else they take the per-lane path)
    const bool xwin = span + 99 <= AO_WIN;
    const float dp = need ? depth[(size_t)y * W + x] : -1.0f;
    // wave w marches direction 2w (an axis, cheap) then 2w+1 (a diagonal,
    // costlier): four waves of equal work, none idling at the barrier
    for (int half = 0; half < 2; half++) {
    const int wu = __builtin_amdgcn_readfirstlane(2 * wv + half);
    const int* axis = reinterpret_cast<const int*>(ao_tab + AO_TAB_AXIS) + (wu >> 1) * 4;
    // Directions 0 and 4 march along x, the major offset exactly +-t.  The
    // minor one, the float32 cos/sin residue rounded onto the row
    // coordinate, moves the tap to another row from some t on only near the
    // frame's bottom rows (it is below half an ulp of Y from Y = 257 up):
    // tap t is on row y for t < x_t1 and on row x_y2 from there, checked
    // here for this row (x_rows: every tap on one of the two, both in the
    // frame).  Where the chunk's taps fit it (xwin), both rows' depths
    // x_first+1 .. x_last+99 (or x_first-99 .. x_last-1) are staged, 0
    // outside the row; a wider chunk takes the per-lane path.
    int x_t1 = 100, x_y2 = y;
    bool x_rows = false;
    if (ao_fast && (wu & 3) == 0 && axis[0] && axis[2] != 0x7fffffff) {
        const float* oy_t = ao_tab + AO_TAB_OFF + wu * 200 + 100;  // fl(dir.y * t)
        const float Yf = (float)y;
        const int t_a = lane, t_b = lane + 64;
        const int iy_a = (int)(Yf + oy_t[t_a]), iy_b = t_b < 100 ? (int)(Yf + oy_t[t_b]) : y;
        const uint64_t off_a = __ballot(t_a >= 1 && iy_a != y), off_b = __ballot(t_b < 100 && iy_b != y);
        x_t1 = off_a ? __ffsll((unsigned long long)off_a) - 1 : (off_b ? 64 + __ffsll((unsigned long long)off_b) - 1 : 100);
        x_y2 = x_t1 < 64 ? __shfl(iy_a, x_t1) : (x_t1 < 100 ? __shfl(iy_b, x_t1 - 64) : y);
        x_t1 = __builtin_amdgcn_readfirstlane(x_t1), x_y2 = __builtin_amdgcn_readfirstlane(x_y2);
        const bool bad = (t_a >= x_t1 && iy_a != x_y2) | (t_b >= x_t1 && t_b < 100 && iy_b != x_y2);
        x_rows = __ballot(bad) == 0 && (unsigned)x_y2 < (unsigned)H;
        if (x_rows && xwin) {
            const int ws = axis[1] > 0 ? x_first + 1 : x_first - 99;
            for (int r = 0; r < (x_t1 < 100 ? 2 : 1); r++) {
                const size_t row = (size_t)(r ? x_y2 : y) * W;
                for (int j = lane; j < span + 99; j += 64) {
                    const int px = ws + j;
                    s_win[wu >> 2][r][j] = (unsigned)px < (unsigned)W ? depth[row + px] : 0.0f;
                }
            }
        }
    }
    // Diagonal marches: where the chunk's first and last pixels (so every
    // lane) put tap t's x in one binade, fl(X + fl(dir.x * t)) - X is the
    // same for the whole chunk (the table's invariant), so every lane's tap
    // is the first pixel's shifted by the lane's column, at one distance:
    // offset and 1/distance per t, and a mask of those t, for the march
    // below.
    uint64_t uni_lo = 0, uni_hi = 0;
    constexpr int ROW_Y = 100 * 8, ROW_X = AO_EXP_N * ROW_Y, BIAS = -AO_EXP_MIN * (ROW_X + ROW_Y);
    if (ao_fast && (wu & 1)) {
        const int k = wu >> 1;
        const float* ox_t = ao_tab + AO_TAB_OFF + wu * 200;
        const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(ao_tab + AO_TAB_DIST + 2 * k * AO_DIST_PER_DIR), (short)0, AO_DIST_PER_DIR * 8, BUF_DWORD3);
        const float X0 = (float)x_first, X63 = (float)x_last, Yf = (float)y;
        bool u0 = false, u1 = false;
        for (int h = 0; h < 2; h++) {
            const int t = lane + 64 * h;
            bool uni = false;
            if (t < 100) {
                const float lo = X0 + ox_t[t], hi = X63 + ox_t[t], cy = Yf + ox_t[100 + t];
                const int kl = __builtin_amdgcn_frexp_expf(lo), kh = __builtin_amdgcn_frexp_expf(hi);
                uni = (lo > 0.0f) & (kl == kh);
                const int iy = (int)cy, ky = __builtin_amdgcn_frexp_expf(cy);
                s_dbase[k][t] = (iy * W + (int)lo) * 4;
                s_drow[k][t] = (unsigned)iy < (unsigned)H ? iy * W : -(1 << 28);
                s_dtab[k][t] = ky * ROW_Y + t * 8 + BIAS;
                const int toff = kl * ROW_X + ky * ROW_Y + t * 8 + BIAS;
                s_drinv[k][t] =
                    uni ? __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(trs, toff, 0, 0)) : 0.0;
            }
            (h ? u1 : u0) = uni;
        }
        uni_lo = __ballot(u0);
        uni_hi = __ballot(u1);
    }
    // the window and the diagonal tables above are this wave's own LDS:
    // its writes are visible to its lanes past a wave-level fence
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float m = 0.0f;
    {
        const float dx = c.ao_dir[2 * wu], dy = c.ao_dir[2 * wu + 1];
        const float X = (float)x, Y = (float)y;
        (void)dx;
        // the tap with the largest exact elevation/distance quotient; its
        // rounded quotient is the largest rounded quotient.  The general
        // path compares exact float64 cross products (e * d_best > e_best * d:
        // products of two float32 are exact).  The table paths compare
        // q = e * fl64(1 / d) instead, within 2^-52 of e / d: two distinct
        // quotients of float32 e and d (d a float32 >= 1) differ relatively by
        // at least 2^-47, and e / d lies at least 2^-47 (relatively) from any
        // float32 rounding boundary, so q orders distinct quotients exactly
        // (equal ones may order either way, and round alike) and float(q)
        // is the correctly rounded float32 quotient.
        double e_best = -__builtin_inf(), d_best = 1.0, q_best = -__builtin_inf();
        bool nan = false;
        // depth taps through a buffer descriptor: an out-of-range offset
        // reads 0 instead of faulting, so a masked lane needs no clamped
        // address (its value is never used)
        const __amdgpu_buffer_rsrc_t drs =
            __builtin_amdgcn_make_buffer_rsrc((void*)depth, (short)0, W * H * 4, BUF_DWORD3);
        const float* off_x = ao_tab + AO_TAB_OFF + wu * 200;  // fl(dir.x * t), t = 0..99 (uniform)
        const float* off_y = off_x + 100;                     // fl(dir.y * t)
        // one tap: the running argmax of elevation / distance, exact
        // (bitwise & keeps it a select, not a branch that would serialise
        // the loads of consecutive taps)
        auto take = [&](bool v, float e, double d64) {
            const double e64 = (double)e;
            const bool better = v & (e64 * d_best > e_best * d64);
            e_best = better ? e64 : e_best;
            d_best = better ? d64 : d_best;
            nan |= v & (e != e);
        };
        auto take_q = [&](bool v, float e, double rinv) {
            const double q = (double)e * rinv;
            q_best = (v & (q > q_best)) ? q : q_best;
            nan |= v & (e != e);
        };
        // every depth finite: no quotient is NaN (an infinite e gives an
        // infinite q), so the NaN test can go (uniform)
        const bool finite = cnt->depth_nonfinite == 0;
        const double* rinv_t = reinterpret_cast<const double*>(ao_tab + AO_TAB_T64);  // fl64(1 / t) (uniform)
        // Axis direction, the segment away from the small minor coordinates:
        // the major offset is exactly +-t and the minor one leaves the tap on
        // the pixel's row (column), so tap t is the pixel t steps along it,
        // in the buffer while t < tlim.  Up to the wave's smallest tlim no
        // lane needs the mask (lanes without AO compute values nobody reads):
        // a plain running max of q.
        const bool row_fast = ao_fast && (wu & 1) == 0 && finite && (axis[0] ? y : x_first) >= axis[2];
        if (ao_fast && finite && x_rows && xwin) {  // (directions 0 and 4, whose windows were loaded)
            // along x: the lane at column li reads tap t at window entry
            // li - 1 + t (li + 99 - t) of its row's window: taps [1, x_t1)
            // row y's, the rest row x_y2's.  Up to the wave's smallest room
            // to the border (tmin) no lane needs the mask.
            const int sgn = axis[1];
            const int base = sgn > 0 ? li - 1 : li + 99;
            const int tlim = need ? (sgn > 0 ? W - x : x + 1) : 100;
            int tl = tlim;  // the wave's smallest
            for (int k = 32; k >= 1; k >>= 1) tl = min(tl, __shfl_xor(tl, k));
            const int tmin = min(__builtin_amdgcn_readfirstlane(tl), 100);
            const int n_rows = x_t1 < 100 ? 2 : 1;
#pragma nounroll
            for (int r = 0; r < n_rows; r++) {
                const int t_lo = r ? x_t1 : 1, t_hi = r ? 100 : x_t1;
                const int t_mid = min(max(tmin, t_lo), t_hi);
                const float* win = s_win[wu >> 2][r];
                // one loop per sign, so each tap's window entry is a constant
                // offset from the lane's address
                if (sgn > 0) {
                    const float* wl = win + li - 1;
#pragma unroll 4
                    for (int ti = t_lo; ti < t_mid; ti++)
                        q_best = __builtin_fmax(q_best, (double)(wl[ti] - dp) * rinv_t[ti]);
                } else {
                    const float* wl = win + li + 99;
#pragma unroll 4
                    for (int ti = t_lo; ti < t_mid; ti++)
                        q_best = __builtin_fmax(q_best, (double)(wl[-ti] - dp) * rinv_t[ti]);
                }
                for (int ti = t_mid; ti < t_hi; ti++) {
                    const double q = (double)(win[base + sgn * ti] - dp) * rinv_t[ti];
                    q_best = (need & (ti < tlim) & (q > q_best)) ? q : q_best;
                }
            }
        } else if (row_fast && !axis[0]) {
            // along y: tap t is row y +- t, the same for every lane, in the
            // buffer for t < tl; the march ends there for the whole wave
            const int sgn = axis[1];
            const int tl = min(sgn > 0 ? H - y : y + 1, 100);
            // the row's offset is uniform (the load's scalar offset), the
            // lane's column its vector offset
            const int xo = x * 4;
#pragma unroll 8
            for (int ti = 1; ti < tl; ti++) {
                const int row = (y + sgn * ti) * W * 4;
                const float e = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(drs, xo, row, 0)) - dp;
                q_best = __builtin_fmax(q_best, (double)e * rinv_t[ti]);
            }
        } else if (ao_fast && (wu & 1) == 0) {
            // Axis directions (a = 0, pi/2, pi, 3pi/2 in float32): one offset
            // is exactly +-t, the other the tiny cos/sin residue (<= 8.8e-8 t)
            // rounded onto the coordinate grid, under 2e-5 for coordinates
            // below 4096.  Its square (< 4e-10) is under half an ulp of t^2
            // (>= 6e-8), so Len = sqrt(t^2) = t exactly (also checked over
            // every coordinate pair below 4096 and every t).
#pragma unroll 3
            for (int ti = 1; ti < 100; ti++) {
                const float cx = X + off_x[ti], cy = Y + off_y[ti];
                const int ix = (int)cx, iy = (int)cy;
                const bool v = need & ((unsigned)ix < (unsigned)W) & ((unsigned)iy < (unsigned)H);
                const uint32_t off = (mul_u24((unsigned)iy, (unsigned)W) + (unsigned)ix) * 4u;
                take_q(v, __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(drs, (int)off, 0, 0)) - dp, rinv_t[ti]);
            }
        } else if (ao_fast) {
            // Diagonal directions: X - fl(X + fl(dir.x * t)) is fixed by t and
            // the binade of the tap coordinate (checked when the table is
            // built, for every coordinate below 4096), so the distance
            // Len = sqrt32(fma64(ex, ex, ey * ey)) is fixed by (binade of cx,
            // binade of cy, t); the table holds fl64(1 / Len).
            const double* dtab =
                reinterpret_cast<const double*>(ao_tab + AO_TAB_DIST) + ((wu - 1) >> 1) * AO_DIST_PER_DIR;
            const __amdgpu_buffer_rsrc_t trs =
                __builtin_amdgcn_make_buffer_rsrc((void*)dtab, (short)0, AO_DIST_PER_DIR * 8, BUF_DWORD3);
            // byte offset of (kx, ky, t) = kx * ROW_X + ky * ROW_Y + t * 8 + BIAS
            struct Tap {
                double rinv;
                float e;
                bool in;  // in the buffer
            };
            const int kd = wu >> 1;
            auto tap = [&](int ti) {
                const float cx = X + off_x[ti];
                const int ix = (int)cx;
                // an exponent outside the table only occurs for a tap outside
                // the buffer (checked when the table is built); its offset
                // then reads 0 or another entry, masked either way.  The row
                // (iy = (int)(Y + fl(dir.y * t)), its exponent) is the chunk's,
                // from the tables above.
                const int kx = __builtin_amdgcn_frexp_expf(cx);
                const int row = s_drow[kd][ti];
                const int toff = mul_i24(kx, ROW_X) + s_dtab[kd][ti];
                Tap r;
                r.rinv = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(trs, toff, 0, 0));
                r.in = ((unsigned)ix < (unsigned)W) & (row >= 0);
                const uint32_t off = ((unsigned)row + (unsigned)ix) * 4u;
                r.e = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(drs, (int)off, 0, 0)) - dp;
                return r;
            };
            {
                const Tap a = tap(1);
                take_q(need & a.in & (a.rinv <= 1.0), a.e, a.rinv);  // distance < 1 only occurs at t = 1
            }
            int ti = 2;
            if (finite) {
                // Taps with t <= r / (|dir| * 1.001), r the room to the border
                // the direction heads for, stay in the buffer whatever the
                // rounding of fl(X + fl(dir * t)) (off by far less than the
                // 0.1 % margin): up to the wave's smallest such bound no lane
                // needs the mask.
                // And from t >= (r + 2) / (|dir| * 0.999) + 2 on, a lane's tap
                // lies more than a pixel past the border (far beyond the
                // rounding), where its march has ended: past the wave's
                // largest such bound no lane has a tap left.
                int tl = 100, tu = 0;
                if (need) {
                    const float rx = dx > 0 ? (float)(W - 1 - x) : X, ry = dy > 0 ? (float)(H - 1 - y) : Y;
                    const float ax = __builtin_fabsf(dx), ay = __builtin_fabsf(dy);
                    const float lim = fminf(rx / (ax * 1.001f), ry / (ay * 1.001f));
                    tl = min((int)lim + 1, 100);
                    const float out = fminf((rx + 2.0f) / (ax * 0.999f), (ry + 2.0f) / (ay * 0.999f));
                    tu = min((int)out + 2, 100);
                }
                for (int k = 32; k >= 1; k >>= 1) tl = min(tl, __shfl_xor(tl, k)), tu = max(tu, __shfl_xor(tu, k));
                const int tmin = __builtin_amdgcn_readfirstlane(tl), tmax = __builtin_amdgcn_readfirstlane(tu);
                const uint32_t lane4 = (uint32_t)li * 4u;
                for (; ti + 7 < tmin; ti += 8) {
                    const uint64_t bits = ti < 64 ? (uni_lo >> ti) | (ti > 56 ? uni_hi << (64 - ti) : 0ull)
                                                  : uni_hi >> (ti - 64);
                    if ((bits & 0xFFull) == 0xFFull) {  // eight taps of one shape across the wave
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const uint32_t off = (uint32_t)s_dbase[kd][ti + j] + lane4;
                            const float e =
                                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(drs, (int)off, 0, 0)) - dp;
                            q_best = __builtin_fmax(q_best, (double)e * s_drinv[kd][ti + j]);
                        }
                    } else {
#pragma unroll 4
                        for (int j = 0; j < 8; j++) {
                            const Tap a = tap(ti + j);
                            q_best = __builtin_fmax(q_best, (double)a.e * a.rinv);
                        }
                    }
                }
                for (; ti < tmin; ti++) {
                    const Tap a = tap(ti);
                    q_best = __builtin_fmax(q_best, (double)a.e * a.rinv);
                }
#pragma unroll 2
                for (; ti < tmax; ti++) {
                    const Tap a = tap(ti);
                    const double q = (double)a.e * a.rinv;
                    q_best = (need & a.in & (q > q_best)) ? q : q_best;
                }
            } else {
#pragma unroll 2
                for (; ti < 100; ti++) {
                    const Tap a = tap(ti);
                    take_q(need & a.in, a.e, a.rinv);
                }
            }
        } else {
            float t = 1.0f;  // t = 0 is the pixel itself: distance 0 < 1, skipped
#pragma unroll 3
            for (int ti = 1; ti < 100; ti++, t += 1.0f) {
                const float cx = X + dx * t, cy = Y + dy * t;
                const int ix = (int)cx, iy = (int)cy;  // |cx|, |cy| < 2^31: Go's int() truncation
                const float ex = X - cx, ey = Y - cy;
                // Len of (ex, ey, 0, 0): FMA(ex,ex, FMA(ey,ey, FMA(0,0,0*0))) where
                // the inner FMA(ey,ey,+0) is exactly the float32 product ey*ey
                const float dist = sqrt32_normal(fma64(ex, ex, ey * ey));
                const bool v = need & ((unsigned)ix < (unsigned)W) & ((unsigned)iy < (unsigned)H) & (dist >= 1.0f);
                const uint32_t off = (mul_u24((unsigned)iy, (unsigned)W) + (unsigned)ix) * 4u;
                take(v, __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(drs, (int)off, 0, 0)) - dp, (double)dist);
            }
        }
        const float maxr = ao_fast ? (float)q_best : (float)e_best / (float)d_best;
        m = nan ? __builtin_nanf("") : (maxr > 0 ? (float)atan64((double)maxr) : 0.0f);
    }
    s_m[wu][lane] = m;
    }
    __syncthreads();
    if (wv == 0 && need) {
        float total = 0.0f;
        for (int k = 0; k < 8; k++) total += 1.57079632679489661923f - s_m[k][lane];
        total /= 12.5663706143591729539f;
        ao[(size_t)y * W + x] = (float)pow64((double)total, 10000.0);
    }
}
__global__ __launch_bounds__(256) void k_ao(k_ao_args a) {
    k_ao_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_ao_wide(const k_ao_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_ao_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_ao_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// The colour of every uncovered pixel (render/raster.go:324-328 with
// raster_screen.go:75-87): DrawFragment hands `shade` the zero fragment, whose
// X = Y = 0 and MaterialID = 0, so it is shaded from pixel (0,0)'s stored
// fragment with material table[0]; if pixel (0,0) is uncovered it is the
// Background.  With no material 0 the colour is pixel (0,0)'s own final one.
struct k_shade_bg_args {
    DevScene s;
    ShadeConsts c;
    ShadeBuffers b;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_shade_bg_body(const k_shade_bg_args& a_, const VB vb) {
    const DevScene& s = a_.s;
    const ShadeConsts& c = a_.c;
    const ShadeBuffers& b = a_.b;

    if (threadIdx.x != 0 || vb.bx != 0) return;
    // the frame's counters are final here: hand them to the host's pinned
    // copy (main then shadow), in place of a copy after the frame
    if (b.cnt_host)
        for (int i = 0; i < (int)(2 * sizeof(Counters) / 4); i++)
            reinterpret_cast<int32_t*>(b.cnt_host)[i] = reinterpret_cast<const int32_t*>(b.cnt)[i];
    if (b.cnt->overflow) return;
    const uint64_t k00 = b.vis[0];
    float lin[3] = {__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf("")};
    uint32_t col = c.background;
    if (k00 != VIS_EMPTY) {
        GFrag f = fetch_fragment(s, c, b, 0, 0, k00);
        bool hl;
        float own_lin[3];
        uint32_t own = shade_fragment(s, c, b, f, 0, 0, f.mat, f.col, own_lin, hl);
        if (s.n_materials > 0) {
            col = shade_fragment(s, c, b, f, 0, 0, 0, f.col, lin, hl);
            if (!hl) lin[0] = lin[1] = lin[2] = __builtin_nanf("");
        } else {
            col = own;
        }
    }
    b.bg[0] = col;
    b.bg_lin[0] = lin[0], b.bg_lin[1] = lin[1], b.bg_lin[2] = lin[2];
}
__global__ __launch_bounds__(64) void k_shade_bg(k_shade_bg_args a) {
    k_shade_bg_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(64) void k_shade_bg_wide(const k_shade_bg_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_shade_bg_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_shade_bg_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// Deferred pass (DrawFragments(shade) then DrawFragments(GammaCorrection),
// render/raster.go:300-314,369-376): one thread per pixel, each wave an 8x8
// block of a 16x16 tile.
struct k_shade_args {
    DevScene s;
    ShadeConsts c;
    ShadeBuffers b;
    int32_t gx, gy;  // the grid of this launch
};
// One 16x16 tile of the shading: a pixel per thread, four 8x8 waves.
__device__ __forceinline__ void shade_tile(const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b, int tile) {
    const int tiles_x = c.v.tiles_x;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int x = (tile % tiles_x) * TILE + (wv & 1) * 8 + (lane & 7);
    const int y = (tile / tiles_x) * TILE + (wv >> 1) * 8 + (lane >> 3);
    if (x >= c.v.W || y >= c.v.H) return;
    const size_t o = (size_t)y * c.v.W + x;
    const uint64_t key = b.vis[o];
    uint32_t col;
    float lin[3] = {__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf("")};
    bool hl = false;
    uint8_t flag = 0;
    if (key == VIS_EMPTY) {
        if (b.col_pre) {  // (the second half shades uncovered pixels)
            b.post_flag[o] = 2;
            return;
        }
        col = b.bg[0];
        lin[0] = b.bg_lin[0], lin[1] = b.bg_lin[1], lin[2] = b.bg_lin[2];
        hl = true;
    } else {
        GFrag f = fetch_fragment(s, c, b, x, y, key);
        col = shade_fragment(s, c, b, f, x, y, f.mat, f.col, lin, hl);
        if (b.col_pre) flag = f.mat >= 0 && s.mats[f.mat].ambient_occlusion;
    }
    if (b.col_pre) {
        // first half of a split shading: the colour before AO, whether the
        // AO factor is still to come, and the linear sums (which AO does
        // not touch)
        b.col_pre[o] = col;
        b.post_flag[o] = flag;
        if (b.linear) {
            b.linear[3 * o] = hl ? lin[0] : __builtin_nanf("");
            b.linear[3 * o + 1] = hl ? lin[1] : __builtin_nanf("");
            b.linear[3 * o + 2] = hl ? lin[2] : __builtin_nanf("");
        }
        return;
    }
    if (b.pregamma) b.pregamma[o] = col;
    if (b.linear) {
        b.linear[3 * o] = hl ? lin[0] : __builtin_nanf("");
        b.linear[3 * o + 1] = hl ? lin[1] : __builtin_nanf("");
        b.linear[3 * o + 2] = hl ? lin[2] : __builtin_nanf("");
    }
    b.image[(size_t)(c.v.H - 1 - y) * c.v.W + x] = gamma(c, col);
}
__device__ __forceinline__ void k_shade_body(const k_shade_args& a_, const VB vb) {
    const DevScene& s = a_.s;
    const ShadeConsts& c = a_.c;
    const ShadeBuffers& b = a_.b;

    if (b.cnt->overflow) return;
    const int tiles_x = c.v.tiles_x, tiles_y = c.v.tiles_y;
    // workgroup i shades tile i: the dispatcher deals consecutive
    // workgroups round-robin to the 8 XCDs, so every XCD gets a share of
    // every part of the frame (the covered pixels, whose shading is the
    // work, crowd where the ground is)
    const int tile = vb.bx;
    if (tile < tiles_x * tiles_y) shade_tile(s, c, b, tile);
}
__global__ __launch_bounds__(256) void k_shade(k_shade_args a) {
    k_shade_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_shade_wide(const k_shade_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_shade_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_shade_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// The fixed shadow visibility of the differentiable shading: per pixel of the
// last frame, the number k_shade halves the colour by -- the lights for which
// shadingVisibility reports true (render/shadow.go:230-283, summed as
// render/raster.go:339-347 sums it) at a covered pixel whose material
// receives shadows; 0 elsewhere.
__global__ __launch_bounds__(256) void k_shadow_count(DevScene s, ShadeConsts c, ShadeBuffers b,
                                                      uint8_t* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)c.v.W * c.v.H) return;
    const int x = (int)(i % (size_t)c.v.W), y = (int)(i / (size_t)c.v.W);
    const uint64_t key = b.vis[i];
    int occ = 0;
    if (key != VIS_EMPTY && c.shadow_map) {
        const GFrag f = fetch_fragment(s, c, b, x, y, key);
        if (f.mat >= 0 && s.mats[f.mat].receive_shadow) occ = occluded(s, c, b, x, y, f.depth);
    }
    out[i] = (uint8_t)occ;
}

// Second half of a split shading (the first half, k_shade with col_pre set,
// ran beside k_ao): the AO factor on the covered pixels whose material has
// AO (AmbientOcclusionShade after the shadow halving, raster.go:339-357),
// the uncovered pixels' colour, then pre-gamma, gamma and the y-flipped store
// -- what k_shade does after shade_fragment.
__global__ __launch_bounds__(256) void k_shade_post(DevScene s, ShadeConsts c, ShadeBuffers b) {
    const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (o >= (size_t)c.v.W * c.v.H || b.cnt->overflow) return;
    const int x = (int)(o % (size_t)c.v.W), y = (int)(o / (size_t)c.v.W);
    const uint8_t flag = b.post_flag[o];  // (the first half's: no dependent loads here)
    uint32_t col;
    if (flag == 2) {
        col = b.bg[0];
        if (b.linear) b.linear[3 * o] = b.bg_lin[0], b.linear[3 * o + 1] = b.bg_lin[1], b.linear[3 * o + 2] = b.bg_lin[2];
    } else {
        col = b.col_pre[o];
        if (flag) col = ao_apply(col, b.ao[o]);
    }
    if (b.pregamma) b.pregamma[o] = col;
    b.image[(size_t)(c.v.H - 1 - y) * c.v.W + x] = gamma(c, col);
}

// ------------------------------------------------------------- launchers
static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------- resize
// resizeRGBA (internal/imageutil/resize.go:76-118), one output pixel per
// thread: out[y][x] filters row x of `in` (rows of in_w RGBA8 pixels) at
// output position y with createWeights8's 8-bit weights.  Two launches make
// imageutil.Resize: the first writes its result transposed.
struct k_resize_args {
    const uint32_t* in;
    int32_t in_w;
    uint32_t* out;
    int32_t out_w;
    int32_t out_rows;
    const int16_t* coeffs;
    const int32_t* start;
    int32_t fl;
    int32_t gx, gy;  // the grid of this launch
};
__device__ __forceinline__ void k_resize_body(const k_resize_args& a_, const VB vb) {
    const uint32_t* __restrict__ in = a_.in;
    const int32_t in_w = a_.in_w;
    uint32_t* __restrict__ out = a_.out;
    const int32_t out_w = a_.out_w;
    const int32_t out_rows = a_.out_rows;
    const int16_t* __restrict__ coeffs = a_.coeffs;
    const int32_t* __restrict__ start = a_.start;
    const int32_t fl = a_.fl;

    const int x = vb.bx * blockDim.x + threadIdx.x, y = vb.by;
    if (x >= out_w || y >= out_rows) return;
    const uint32_t* row = in + (size_t)x * in_w;
    const int maxX = in_w - 1;
    int32_t acc[4] = {0, 0, 0, 0}, sum = 0;
    for (int i = 0; i < fl; i++) {
        const int32_t coeff = coeffs[y * fl + i];
        if (coeff == 0) continue;
        int xi = start[y] + i;
        xi = (unsigned)xi < (unsigned)maxX ? xi : (xi >= maxX ? maxX : 0);
        const uint32_t p = row[xi];
        for (int k = 0; k < 4; k++) acc[k] += coeff * (int32_t)((p >> (8 * k)) & 0xffu);
        sum += coeff;
    }
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) {
        const int v = acc[k] / sum;  // truncating division, as Go's
        o |= (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)) << (8 * k);
    }
    out[(size_t)y * out_w + x] = o;
}
__global__ __launch_bounds__(256) void k_resize(k_resize_args a) {
    k_resize_body(a, VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y});
}
__global__ __launch_bounds__(256) void k_resize_wide(const k_resize_args* __restrict__ a, const int32_t* __restrict__ start,
        int32_t n) {
    const int v = wide_view(start, n, (int)blockIdx.x);
    const k_resize_args& x = a[v];
    const int local = (int)blockIdx.x - start[v];
    if (local >= x.gx * x.gy) return;  // padding to a multiple of 8 workgroups
    k_resize_body(x, VB{local % x.gx, local / x.gx, x.gx, x.gy});
}

// ------------------------------------------------------------- wide launches
// While a WideBatch records (wide_record), every launch_* call appends its
// argument block and grid to the group of its stage instead of launching;
// wide_launch then issues one *_wide launch per group, whose workgroups find
// their view's arguments by a binary search over the group's block offsets
// (each view's range padded to a multiple of 8 so that its workgroups keep
// the XCD spread of its own launch).  Stage numbers come from the caller
// (wide_stage), in the order a frame issues them, so views that skip a stage
// simply have no entry in its group.
struct WideGroup {
    const void* kernel = nullptr;  // the single-view kernel of every entry
    bool mixed = false;
    int bs = 0;
    std::vector<uint8_t> args;
    std::vector<int32_t> start{0};
    std::function<void(hipStream_t, const uint8_t*, const int32_t*, int32_t, int32_t)> launch;
    size_t off_args = 0, off_start = 0;
};
struct WideBatch {
    std::map<int, WideGroup> groups;
    int stage = 0, sub = 0;
};
static thread_local WideBatch* g_wide = nullptr;  // the calling thread's recording batch

WideBatch* wide_new() { return new WideBatch(); }
void wide_free(WideBatch* w) { delete w; }
void wide_record(WideBatch* w) { g_wide = w; }
void wide_stage(int stage) {
    if (g_wide) g_wide->stage = stage, g_wide->sub = 0;
}
size_t wide_layout(WideBatch* w) {
    size_t off = 0;
    for (auto& kv : w->groups) {
        WideGroup& g = kv.second;
        if (g.mixed) return 0;
        g.off_args = off;
        off += (g.args.size() + 255) & ~(size_t)255;
        g.off_start = off;
        off += (g.start.size() * 4 + 255) & ~(size_t)255;
    }
    return off;
}
int wide_upload(const WideBatch* w, uint8_t* dev) {
    for (const auto& kv : w->groups) {
        const WideGroup& g = kv.second;
        if (hipMemcpy(dev + g.off_args, g.args.data(), g.args.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dev + g.off_start, g.start.data(), g.start.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return -1;
    }
    return 0;
}
void wide_launch(const WideBatch* w, hipStream_t st, const uint8_t* dev) {
    for (const auto& kv : w->groups) {
        const WideGroup& g = kv.second;
        const int32_t n = (int32_t)g.start.size() - 1;
        if (n > 0 && g.start.back() > 0)
            g.launch(st, dev + g.off_args, reinterpret_cast<const int32_t*>(dev + g.off_start), n, g.start.back());
    }
}

// One launch of `single` over `grid`, or its entry in the recording batch.
template <class A>
static void go(void (*single)(A), void (*wide)(const A*, const int32_t*, int32_t), dim3 grid, int bs,
               hipStream_t st, A a) {
    a.gx = (int32_t)grid.x, a.gy = (int32_t)grid.y;
    if (!g_wide) {
        hipLaunchKernelGGL(single, grid, dim3(bs), 0, st, a);
        return;
    }
    WideGroup& g = g_wide->groups[g_wide->stage * 8 + g_wide->sub++];
    if (!g.kernel) {
        g.kernel = (const void*)single, g.bs = bs;
        g.launch = [wide, bs](hipStream_t s, const uint8_t* args, const int32_t* start, int32_t n, int32_t total) {
            hipLaunchKernelGGL(wide, dim3(total), dim3(bs), 0, s, reinterpret_cast<const A*>(args), start, n);
        };
    } else if (g.kernel != (const void*)single || g.bs != bs) {
        g.mixed = true;
    }
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&a);
    g.args.insert(g.args.end(), p, p + sizeof(A));
    const int32_t blocks = (int32_t)(grid.x * grid.y);
    g.start.push_back(g.start.back() + ((blocks + 7) & ~7));
}

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<size_t>(blocks_for(n, 256), 4096)), dim3(256), 0, st, p, n, v);
}
void launch_frame_clear(hipStream_t st, uint32_t* cnt, int32_t n_cnt, uint64_t* vis, size_t n_vis) {
    go(k_frame_clear, k_frame_clear_wide, dim3(std::min<size_t>(blocks_for(std::max<size_t>(n_vis, n_cnt), 256), 4096)),
       256, st, k_frame_clear_args{cnt, n_cnt, vis, n_vis, 0, 0});
}
void launch_fill_u64(hipStream_t st, uint64_t* p, size_t n, uint64_t v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(std::min<size_t>(blocks_for(n, 256), 4096)), dim3(256), 0, st, p, n, v);
}
// The clear on its own where the vertex pass has nothing to launch (a group
// of its own in a recorded batch, beside the views' vertex passes).
static void clear_alone(hipStream_t st, const FrameClear& clr) {
    if (!clr.n_cnt && !clr.n_vis) return;
    if (g_wide) g_wide->sub = 1;
    launch_frame_clear(st, clr.cnt, clr.n_cnt, clr.vis, clr.n_vis);
}
void launch_vertex_main(hipStream_t st, const DevScene& s, const M4* trans, const M4* normal, const ViewConsts& vc,
                        float4* tv_scr, float4* tv_nor, float4* tv_wp, uint32_t* tv_col, float2* tv_uv,
                        const FrameClear& clr) {
    if (!s.n_tverts) return clear_alone(st, clr);
    go(k_vertex_main, k_vertex_main_wide, dim3(blocks_for(s.n_tverts, 256)), 256, st,
       k_vertex_main_args{s, trans, normal, vc, tv_scr, tv_nor, tv_wp, tv_col, tv_uv, clr, 0, 0});
}
void launch_vertex_shadow(hipStream_t st, const DevScene& s, const M4* strans, const ShadowConsts& sc, float3* ts,
                          const FrameClear& clr) {
    if (!s.n_tverts || !sc.n_slots) return clear_alone(st, clr);
    go(k_vertex_shadow, k_vertex_shadow_wide, dim3(blocks_for(s.n_tverts, 256)), 256, st,
       k_vertex_shadow_args{s, strans, sc, ts, clr, 0, 0});
}
int setup_blocks(int32_t n_tris) { return (n_tris + 255) / 256; }
int clip_grid(const Capacities& cap) { return (int)std::min<size_t>(blocks_for(cap.clip_tris, CLIP_WAVES), CLIP_GRID); }
void launch_setup_main(hipStream_t st, const DevScene& s, const ViewConsts& vc, const float4* tv_scr,
                       const float4* tv_nor, Prim* prims, uint2* pbox, ClipRec* clips, int32_t* tri_clip_slot,
                       ClipJob* clip_list, int32_t* tile_count, int32_t* blk_prims, Counters* cnt, Capacities cap,
                       unsigned long long* vis, int32_t draw_small, const int4* tri_rec) {
    if (!s.n_tris) return;
    go(k_setup_main, k_setup_main_wide, dim3(setup_blocks(s.n_tris)), 256, st,
       k_setup_main_args{s, vc, tv_scr, prims, pbox, clip_list, tile_count, blk_prims, cnt, cap, vis, draw_small,
                         tri_rec, 0, 0});
    go(k_clip, k_clip_wide, dim3(clip_grid(cap)), CLIP_WAVES * 64, st,
       k_clip_args{s, vc, tv_scr, tv_nor, clip_list, prims, pbox, clips, tri_clip_slot, tile_count,
                   blk_prims + setup_blocks(s.n_tris), cnt, cap, 0, 0});
}
void launch_setup_shadow(hipStream_t st, const DevScene& s, const ShadowConsts& sc, const int4* tri_rec,
                         const float3* ts, Prim* prims, uint2* pbox, int32_t* tile_count, Counters* cnt, Capacities cap,
                         uint32_t* smaps) {
    if (!s.n_tris || !sc.n_slots) return;
    go(k_setup_shadow, k_setup_shadow_wide, dim3(blocks_for(s.n_tris, 256), sc.n_slots), 256, st,
       k_setup_shadow_args{s, sc, tri_rec, ts, prims, pbox, tile_count, cnt, cap, smaps, 0, 0});
}
// Each global triangle's record, fixed by the scene (k_tri_rec).
__global__ __launch_bounds__(256) void k_tri_rec(DevScene s, int4* tri_rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= s.n_tris) return;
    const int o = s.tri_obj[g];
    const int mesh = s.obj_mesh[o];
    const int t = s.mesh_tri_off[mesh] + (g - s.obj_tri_start[o]);
    if (!s.tri_valid[t]) {
        tri_rec[g] = make_int4(-1, -1, -1, -1);
        return;
    }
    const int vb = s.obj_vtx_start[o];
    // the material in the flat table (-1: vertex colour), matAt's rule
    int mat = s.tri_mat[t];
    if (mat >= 0) mat += s.obj_mat_base[o];
    tri_rec[g] = make_int4(vb + s.tri_v[3 * t], vb + s.tri_v[3 * t + 1], vb + s.tri_v[3 * t + 2],
                           (mat >= 0 && mat < s.n_materials) ? mat : -1);
}
void launch_tri_rec(hipStream_t st, const DevScene& s, int4* tri_rec) {
    if (s.n_tris) hipLaunchKernelGGL(k_tri_rec, dim3(blocks_for(s.n_tris, 256)), dim3(256), 0, st, s, tri_rec);
}
int scan_blocks(int32_t n) { return std::max(1, (n + SCAN_BLOCK - 1) / SCAN_BLOCK); }
void launch_scan(hipStream_t st, int32_t* counts, int32_t n, int32_t* offsets, int32_t* chunk_off, int32_t* cursor,
                 int32_t* bsum, int32_t* item_tile, int32_t cap_items, const int32_t* blk_prims, int32_t n_blk,
                 Counters* cnt, int32_t item_prims) {
    const int nb = n <= SCAN_WHOLE_MAX ? 1 : scan_blocks(n);
    // (the sums stage always records, so the apply stage keeps its place)
    if (g_wide) g_wide->sub = 0;
    if (nb > 1)
        go(k_scan_sums, k_scan_sums_wide, dim3(nb), 1024, st, k_scan_sums_args{counts, n, bsum, item_prims, 0, 0});
    if (g_wide) g_wide->sub = 1;
    go(k_scan_apply, k_scan_apply_wide, dim3(nb), 1024, st,
       k_scan_apply_args{counts, n, bsum, offsets, chunk_off, cursor, item_tile, cap_items, blk_prims, n_blk,
                         cnt, item_prims, 0, 0});
}
void launch_bin(hipStream_t st, const uint2* pbox, Counters* cnt, int32_t tiles_x, int32_t shift, int32_t tiles_per_view,
                int32_t shadow, int32_t* cursor, int32_t* list, Capacities cap, int32_t n_fixed, int32_t n_tris) {
    const int max_prims = n_fixed + (shadow ? 0 : cap.clip_subs);
    if (max_prims <= 0) return;
    go(k_bin, k_bin_wide, dim3(blocks_for(max_prims, 256)), 256, st,
       k_bin_args{pbox, cnt, tiles_x, shift, tiles_per_view, shadow, cursor, list, cap, n_fixed, n_tris, 0, 0});
}
template <int MODE>
static void launch_raster_t(hipStream_t st, int tile, int n_workgroups, const k_raster_args& a) {
    // (in a recorded batch each tile side is a group of its own in the
    // stage, so views with different tiles still share wide launches)
    if (g_wide) g_wide->sub = tile == 8 ? 0 : tile == 16 ? 1 : 2;
    if (tile == 8) go(k_raster<MODE, 8>, k_raster_wide<MODE, 8>, dim3(n_workgroups), 256, st, a);
    else if (tile == 32) go(k_raster<MODE, 32>, k_raster_wide<MODE, 32>, dim3(n_workgroups), 256, st, a);
    else go(k_raster<MODE, 16>, k_raster_wide<MODE, 16>, dim3(n_workgroups), 256, st, a);
}
void launch_raster(hipStream_t st, int mode, int n_workgroups, const Prim* prims, const int32_t* tile_off,
                   const int32_t* chunk_off, int32_t* item_tile, int32_t cap_items, const int32_t* tile_list,
                   Counters* cnt, const RasterTarget& rt, int32_t item_prims, int32_t tile) {
    const k_raster_args a{prims, tile_off, chunk_off, item_tile, tile_list, cnt, rt, cap_items, item_prims, 0, 0};
    if (mode == MODE_SHADOW) launch_raster_t<MODE_SHADOW>(st, tile, n_workgroups, a);
    else if (mode == MODE_EXACT) launch_raster_t<MODE_EXACT>(st, tile, n_workgroups, a);
    else launch_raster_t<MODE_VIS>(st, tile, n_workgroups, a);
}
void launch_resolve(hipStream_t st, bool exact, const RasterTarget& rt, Counters* cnt, float* depth, const DevScene& s,
                    int32_t* ao_xs, int32_t* ao_cnt, const int4* tri_rec) {
    if (!rt.W || !rt.H) return;
    const k_resolve_args a{rt, cnt, depth, s, ao_xs, ao_cnt, tri_rec, 0, 0};
    if (exact) go(k_resolve<true>, k_resolve_wide<true>, dim3(rt.H), 256, st, a);
    else go(k_resolve<false>, k_resolve_wide<false>, dim3(rt.H), 256, st, a);
}
void launch_ao(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b) {
    const int nb = ((c.v.W + 63) / 64) * c.v.H;
    go(k_ao, k_ao_wide, dim3(((nb + 7) / 8) * 8), 256, st,
       k_ao_args{c, b.ao_xs, b.ao_cnt, b.depth, b.ao, b.ao_tab, b.ao_fast, b.cnt, 0, 0});
}
void launch_shade_bg(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b) {
    go(k_shade_bg, k_shade_bg_wide, dim3(1), 64, st, k_shade_bg_args{s, c, b, 0, 0});
}
void launch_shade(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b) {
    int n_tiles = c.v.tiles_x * c.v.tiles_y;
    go(k_shade, k_shade_wide, dim3((n_tiles + 7) / 8 * 8), 256, st, k_shade_args{s, c, b, 0, 0});
}

bool launch_shade_post(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b) {
    const size_t P = (size_t)c.v.W * c.v.H;
    if (!b.col_pre || !b.post_flag || !b.ao) return false;  // it reads the first half's buffers: nothing launched
    if (P) hipLaunchKernelGGL(k_shade_post, dim3(blocks_for(P, 256)), dim3(256), 0, st, s, c, b);
    return true;
}
void launch_shadow_count(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b,
                         uint8_t* out) {
    const size_t P = (size_t)c.v.W * c.v.H;
    if (P) hipLaunchKernelGGL(k_shadow_count, dim3(blocks_for(P, 256)), dim3(256), 0, st, s, c, b, out);
}
void launch_resize(hipStream_t st, const uint32_t* in, int32_t in_w, uint32_t* out, int32_t out_w, int32_t out_rows,
                   const int16_t* coeffs, const int32_t* start, int32_t fl) {
    if (out_w <= 0 || out_rows <= 0) return;
    go(k_resize, k_resize_wide, dim3(blocks_for(out_w, 256), out_rows), 256, st,
       k_resize_args{in, in_w, out, out_w, out_rows, coeffs, start, fl, 0, 0});
}

}  // namespace prh
