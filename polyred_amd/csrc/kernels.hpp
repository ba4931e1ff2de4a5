// kernels.hpp — host-side launchers of libprh's gfx950 kernels.
#pragma once

#include "prh_internal.hpp"

namespace prh {

struct ShadeConsts {
    ViewConsts v;
    uint8_t gamma_lut[256];   // shader.GammaCorrection per 8-bit value
    uint32_t background;      // render.Background RGBA8
    int32_t shadow_map;       // render.ShadowMap
    int32_t gamma_correct;    // render.GammaCorrection
    int32_t n_lights;
    float ao_dir[16];         // material/ao.go:26 directions {cos, sin} x 8
    int32_t SW, SH;           // shadow map size
    gm::M4 svp;               // ViewportMatrix(SW, SH)
    int32_t shadow_sparse;    // the light views' last row is (0 0 0 1) and the light projections and svp
                              // have zeros everywhere but the diagonal and the last column
};

// A shadow-casting light's transforms for the shading's shadow test where
// they have the sparse shape (light view's last row (0 0 0 1), projection and
// map viewport diagonal plus last column): the coefficients that multiply a
// coordinate in an fma64 held as float64 (exact), so the test converts none
// of them, and the last-column terms it multiplies in float32 as float.
struct ShadowXf {
    double view[9];   // light view rows 0-2, columns 0-2
    double proj[3];   // light projection diagonal
    double svp[3];    // map viewport diagonal
    float view_c[3];  // light view column 3, rows 0-2
    float proj_c[3];
    float svp_c[3];
    float pad[3];
};

struct ShadeBuffers {
    const uint64_t* vis;
    const float* depth;
    const float4* tv_scr;
    const float4* tv_nor;
    const float4* tv_wp;
    const int4* tri_rec;       // per triangle: transformed vertices and flat material (k_tri_rec)
    const uint32_t* tv_col;    // per transformed vertex: colour
    const float2* tv_uv;       // and UV (scenes with textures)
    const ClipRec* clips;
    const int32_t* tri_clip_slot;
    const uint32_t* smaps;     // [slots][SH][SW] ordered depth bits
    const gm::M4* light_view;  // [n_lights]
    const gm::M4* light_proj;  // [n_lights]
    const ShadowXf* light_xf;  // [n_lights], read where ShadeConsts.shadow_sparse is set
    uint32_t* image;           // [H][W] RGBA8, top row first
    uint32_t* pregamma;        // optional [H][W] raster rows
    float* linear;             // optional [H][W][3]
    uint32_t* bg;              // [2] uncovered colour, pixel (0,0)'s colour
    float* bg_lin;             // [3]
    const Counters* cnt;       // main-pass counters, then the shadow ones: nothing is shaded after an overflow
    Counters* cnt_host;        // pinned host copy of both, written once they are final
    float* ao;                 // [H][W] AO factor of the pixels that need it
    int32_t* ao_xs;            // [H][W] per row, the columns of the pixels k_ao computes, left to right (k_resolve)
    int32_t* ao_cnt;           // [H] and their count
    const float* ao_tab;       // AO march tables (tap offsets, diagonal distances)
    int32_t ao_fast;           // the tables hold for this frame size
    uint32_t* col_pre;         // set: k_shade is the first half of a split shading (colour before AO)
    uint8_t* post_flag;        // and per pixel what the second half does: 0 colour as is, 1 with the AO
                               // factor, 2 uncovered (the uncovered colour)
};

// The differentiable shading pass (diff.hip): parameters, fixed inputs and
// gradient outputs, all in HBM.
constexpr int DIFF_MAX_LIGHTS = 8;
struct DiffBuffers {
    const uint64_t* vis;       // fixed visibility (the last frame)
    const float4* tv_scr;      // screen xy + recipw per transformed vertex
    const float4* tv_nor;      // n.Apply(Normal) per transformed vertex (the default normals)
    const float4* tv_wp;       // back-transformed position per transformed vertex
    const float* albedo;       // [Vt][3]
    const float* normal;       // [Vt][3], or null: tv_nor
    const float* kd;           // [M][3]
    const float* ks;           // [M][3]
    const float* shininess;    // [M]
    const float* light_pos;    // [L][3]
    const float* light_int;    // [L]
    float ambient;             // sum of the ambient intensities
    const uint8_t* shadow_count;  // [H][W] fixed shadow visibility: the shading is scaled by 2^-count; or null
    const float4* gbuf;        // [4][H][W] stored G-buffer (diff_store_gbuffer), or null: rebuilt from vis
    const float* target;       // [H][W][3] raster rows (adjoint)
    float* linear;             // [H][W][3] raster rows (forward), NaN where not shaded
    float *g_albedo, *g_normal, *g_kd, *g_ks, *g_shin, *g_lpos, *g_lint;  // zeroed before the adjoint
    double* loss;
    uint32_t want;             // DIFF_* gradient groups to accumulate
};
// Gradient groups (the bits of prh.h's PRH_FIT_* flags).
constexpr uint32_t DIFF_ALBEDO = 1, DIFF_NORMAL = 2, DIFF_KD = 4, DIFF_KS = 8, DIFF_SHININESS = 16,
                   DIFF_LIGHT_POS = 32, DIFF_LIGHT_INT = 64, DIFF_ALL = 127;
void launch_diff(hipStream_t st, bool backward, const DevScene& s, const ViewConsts& v, const DiffBuffers& b);
// One view of a multi-view launch: its blocks are [block0, block0 + nblocks).
struct DiffView {
    ViewConsts v;
    DiffBuffers b;
    int32_t block0, nblocks;
};
int32_t diff_blocks(const ViewConsts& v);
void launch_diff_views(hipStream_t st, bool backward, const DevScene& s, const DiffView* views, int32_t n_views,
                       int32_t total_blocks);
// The stored G-buffer of one view (world position + material, normal +
// shadow scale, view vector, albedo: 4 float4 planes of H*W).
void launch_diff_store_gbuffer(hipStream_t st, const DevScene& s, const ViewConsts& v, const DiffBuffers& b,
                               float4* gbuf);
// Adam (Kingma & Ba 2015, Algorithm 1) on the parameters p = base * (1 + theta)
// listed in idx; t is the device step counter, loss the step's loss, kept
// in hist[t - 1].
struct AdamConsts {
    float lr, beta1, beta2, eps;
    int32_t n, hist_cap;
};
void launch_adam(hipStream_t st, float* par, const float* grad, const float* base, float* theta, float* m, float* v,
                 const int32_t* idx, const float* lower, AdamConsts a, int32_t* step, const double* loss,
                 double* hist);

// Layout of the AO march tables, in float (4-byte) units.  Binade buckets are
// frexp exponents AO_EXP_MIN .. AO_EXP_MIN + AO_EXP_N - 1.
constexpr int AO_EXP_MIN = -10;
constexpr int AO_EXP_N = 24;
constexpr int AO_DIST_PER_DIR = AO_EXP_N * AO_EXP_N * 100;  // doubles
constexpr int AO_TAB_OFF = 0;           // float [8 directions][2][100 t]: fl(dir * t)
constexpr int AO_TAB_T64 = 8 * 200;     // double [100]: fl64(1 / t)
constexpr int AO_TAB_DIST = AO_TAB_T64 + 200;  // double [4 diagonals][AO_EXP_N][AO_EXP_N][100 t]: fl64(1 / distance)
// int32 [4 axis directions][4]: major axis (1 = x), sign of the major
// offset (+-1, exactly +-t), the smallest minor coordinate from which the
// minor offset never moves the tap off its row/column (INT32_MAX: none)
constexpr int AO_TAB_AXIS = AO_TAB_DIST + 2 * 4 * AO_DIST_PER_DIR;
constexpr int AO_TAB_SIZE = AO_TAB_AXIS + 16;
// k_ao's LDS window of a row for its x marches: a chunk of AO pixels
// spanning up to AO_WIN - 99 columns
constexpr int AO_WIN = 448;
// Third descriptor word of a raw (unstrided, untyped) buffer on gfx9 parts.
constexpr uint32_t BUF_DWORD3 = 0x00020000u;

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v);
void launch_fill_u64(hipStream_t st, uint64_t* p, size_t n, uint64_t v);
void launch_frame_clear(hipStream_t st, uint32_t* cnt, int32_t n_cnt, uint64_t* vis, size_t n_vis);
// A frame's start: counters to zero, the visibility buffer emptied.  Each
// stream's first pass (the vertex passes) does its part before its own
// work, with k_frame_clear only where that pass has nothing to launch.
struct FrameClear {
    uint32_t* cnt;
    int32_t n_cnt;
    uint64_t* vis;
    size_t n_vis;
};
void launch_vertex_main(hipStream_t st, const DevScene& s, const gm::M4* trans, const gm::M4* normal,
                        const ViewConsts& vc, float4* tv_scr, float4* tv_nor, float4* tv_wp, uint32_t* tv_col,
                        float2* tv_uv, const FrameClear& clr);
void launch_vertex_shadow(hipStream_t st, const DevScene& s, const gm::M4* strans, const ShadowConsts& sc,
                          float3* ts, const FrameClear& clr);
// The setup's workgroups (each writes its unclipped primitives to
// blk_prims[its index]) and then the clipper's (blk_prims[setup_blocks +
// its index]: its pieces); blk_prims holds setup_blocks(n_tris) + CLIP_GRID.
constexpr int CLIP_GRID = 512;
int setup_blocks(int32_t n_tris);
int clip_grid(const Capacities& cap);
// pbox: the pixel box of every primitive slot (k_bin's input)
void launch_setup_main(hipStream_t st, const DevScene& s, const ViewConsts& vc, const float4* tv_scr,
                       const float4* tv_nor, Prim* prims, uint2* pbox, ClipRec* clips, int32_t* tri_clip_slot,
                       ClipJob* clip_list, int32_t* tile_count, int32_t* blk_prims, Counters* cnt, Capacities cap,
                       unsigned long long* vis, int32_t draw_small, const int4* tri_rec);
void launch_setup_shadow(hipStream_t st, const DevScene& s, const ShadowConsts& sc, const int4* tri_rec,
                         const float3* ts, Prim* prims, uint2* pbox, int32_t* tile_count, Counters* cnt, Capacities cap,
                         uint32_t* smaps);
// Each global triangle's record, fixed by the scene: its transformed
// vertices and its material in the flat table (-1: vertex colour, matAt's
// rule), or x = -1 when Triangle.IsValid is false.  Once per renderer.
void launch_tri_rec(hipStream_t st, const DevScene& s, int4* tri_rec);
// n_blk > 0: also sum the per-workgroup primitive counts of the setup and
// the clipper (blk_prims) into the frame statistics (the main view)
// bsum: 2 * scan_blocks(n) ints of scratch
int scan_blocks(int32_t n);
// Also lists each work item's tile in item_tile and zeroes the counts.
void launch_scan(hipStream_t st, int32_t* counts, int32_t n, int32_t* offsets, int32_t* chunk_off, int32_t* cursor,
                 int32_t* bsum, int32_t* item_tile, int32_t cap_items, const int32_t* blk_prims, int32_t n_blk,
                 Counters* cnt, int32_t item_prims);
// shift: log2 of the raster tile side
void launch_bin(hipStream_t st, const uint2* pbox, Counters* cnt, int32_t tiles_x, int32_t shift, int32_t tiles_per_view,
                int32_t shadow, int32_t* cursor, int32_t* list, Capacities cap, int32_t n_fixed, int32_t n_tris);
// mode: 0 visibility, 1 visibility in exact order (NaN rule), 2 shadow maps;
// tile: the raster tile side, 8, 16 or 32
void launch_raster(hipStream_t st, int mode, int n_workgroups, const Prim* prims, const int32_t* tile_off,
                   const int32_t* chunk_off, int32_t* item_tile, int32_t cap_items, const int32_t* tile_list,
                   Counters* cnt, const RasterTarget& rt, int32_t item_prims, int32_t tile);
void launch_resolve(hipStream_t st, bool exact, const RasterTarget& rt, Counters* cnt, float* depth, const DevScene& s,
                    int32_t* ao_xs, int32_t* ao_cnt, const int4* tri_rec);
void launch_ao(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b);
void launch_shade_bg(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b);
void launch_shade(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b);
// false (and nothing launched) without the colour before AO, the AO factors or the triangle records
bool launch_shade_post(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b);
void launch_shadow_count(hipStream_t st, const DevScene& s, const ShadeConsts& c, const ShadeBuffers& b,
                         uint8_t* out);

// Wide launches: many views' frames as one launch per stage (kernels.hip).
struct WideBatch;
WideBatch* wide_new();
void wide_free(WideBatch* w);
void wide_record(WideBatch* w);  // launch_* calls record into w (nullptr: launch again)
void wide_stage(int stage);      // the stage of the following launch_* calls
size_t wide_layout(WideBatch* w);  // device bytes of the argument arrays (0: unusable)
int wide_upload(const WideBatch* w, uint8_t* dev);
void wide_launch(const WideBatch* w, hipStream_t st, const uint8_t* dev);
// out[y][x] (out_rows x out_w) = 8-bit bilinear filter of row x of `in` at y
void launch_resize(hipStream_t st, const uint32_t* in, int32_t in_w, uint32_t* out, int32_t out_w, int32_t out_rows,
                   const int16_t* coeffs, const int32_t* start, int32_t fl);

}  // namespace prh
