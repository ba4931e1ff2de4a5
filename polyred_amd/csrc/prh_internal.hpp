// prh_internal.hpp — device data layout shared by libprh's host runtime and
// its gfx950 kernels.  Everything lives in HBM, structure-of-arrays where a
// kernel streams it, fixed-size records where a kernel gathers it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gomath.hpp"

namespace prh {

constexpr int TILE = 16;                 // shading tile side (pixels): one 256-thread workgroup
constexpr int TILE_PX = TILE * TILE;     // 256 pixels = 4 wave64s
// The raster tile (binning and the binned raster) is a tuning knob of its
// own: 8, 16 or 32 pixels a side (prh_tuning.raster_tile).
constexpr int RTILE_MIN = 8, RTILE_MAX = 32;
constexpr uint32_t SUB_DIRECT = 7;       // key sub-index of an unclipped triangle
constexpr uint64_t VIS_EMPTY = ~0ull;    // no fragment
constexpr uint32_t KEY_HI_NAN = 0u;      // NaN-depth winner (exact-order pass only)

// Material as the shading kernel reads it (from prh_material).
struct DevMaterial {
    float diffuse[3];   // float32(Diffuse.R/G/B)
    float specular[3];  // float32(Specular.R/G/B)
    float tex[3];       // float32(texture R/G/B)
    float la[3];        // sum_e Intensity_e * float32(tex)  (blinn_cpu.go:44-52)
    float shininess;
    uint32_t tex_rgba;  // texture RGBA8 (alpha is the shaded alpha)
    int32_t flat_shading, ambient_occlusion, receive_shadow;
    int32_t tex_id;     // index into DevScene.textures, or -1: the uniform colour above
};

// A texture with its mip levels (buffer.NewTexture, buffer/texture.go:35-68)
// in DevScene.tex_data: level i is w[i] x h[i] RGBA8 texels at off[i].
constexpr int TEX_MAX_LEVELS = 16;
struct DevTexture {
    int32_t n, use_mipmap, size, pad;
    int32_t off[TEX_MAX_LEVELS], w[TEX_MAX_LEVELS], h[TEX_MAX_LEVELS];
};

struct DevLight {
    int32_t kind;        // 0 point, 1 directional
    float pos[3];
    float dir[3];
    float intensity;
    float color[3];      // float32(Color.R/G/B)
    int32_t cast_shadow;
    int32_t slot;        // shadow-map slot, -1 if none
};

// Scene geometry in HBM.
struct DevScene {
    const float4* vpos;       // [Vm] model-space positions (mesh vertices)
    const float4* vnor;       // [Vm] model-space normals
    const uint32_t* vcol;     // [Vm] RGBA8
    const int32_t* tri_v;     // [Tm][3] mesh-local vertex ids
    const int32_t* tri_mat;   // [Tm] geometry-local material id
    const uint8_t* tri_valid; // [Tm] Triangle.IsValid (model space, per mesh triangle)
    const int32_t* mesh_vtx_off;
    const int32_t* mesh_tri_off;
    const int32_t* obj_mesh;      // [O]
    const int32_t* obj_mat_base;  // [O]
    const int32_t* obj_tri_start; // [O+1] first global triangle of each object
    const int32_t* obj_vtx_start; // [O+1] first transformed vertex of each object
    const int32_t* tri_obj;       // [Tg] object of each global triangle
    const int32_t* vtx_obj;       // [Vt] object of each transformed vertex
    int32_t n_objects;
    int32_t n_tris;               // Tg = sum over objects of mesh triangles
    int32_t n_tverts;             // Vt = sum over objects of mesh vertices
    int32_t n_materials;
    const DevMaterial* mats;
    int32_t n_lights;
    const DevLight* lights;
    const float2* vuv;            // [Vm] texture coordinates (mesh vertices)
    int32_t n_textures;
    const DevTexture* textures;
    const uint32_t* tex_data;     // every level of every texture, RGBA8
    int32_t n_ambient;
    const float* ambient;         // ambient intensities in Scene order
};

// Per-view constants of the main camera (passed by value).
struct ViewConsts {
    gm::M4 vp, vpinv, pinv, vinv, s2w;
    float cam[3];
    int32_t W, H, tiles_x, tiles_y;  // shading tiles (TILE)
    int32_t rtiles_x, rt_shift;      // raster tiles: per row, log2 of the side
    int32_t perspect;
    float vpbox_max_x, vpbox_max_y;  // viewport AABB max (MSAA*W, MSAA*H)
};

// Per-frame shadow constants (all shadow-casting lights at once).
struct ShadowConsts {
    gm::M4 vp;                 // ViewportMatrix(SW, SH)
    int32_t SW, SH, tiles_x, tiles_y;  // raster tiles of one map
    int32_t n_slots, tiles_per_map, rt_shift;
    float vpbox_max_x, vpbox_max_y;  // cullViewFrustum's box: (SW*MSAA, SH*MSAA) (render/cull.go:16-22)
    int32_t direct_max;              // triangles of at most this many texel centres are drawn by the setup
};

// A rasterisable primitive: a whole triangle (sub = 7) or one clipped fan
// piece.  p2.y holds the order key (main view) or the shadow slot.
struct Prim {
    float4 p0;  // ax ay az bx
    float4 p1;  // by bz cx cy
    float4 p2;  // cz key bbox_lo bbox_hi (bit patterns)
};

// Vertex data of one clipped fan piece (render/clipping.go:67-157).
struct ClipRec {
    float4 pos[3];  // screen x y z, parent recipw
    float4 nor[3];  // normal x y z, colour RGBA8 bits
    float4 wp[3];   // back-transformed m_i x y z (render/raster.go:467-469)
    float2 uv[3];   // texture coordinates (render/clipping.go:88-91)
};

// A triangle queued for the clipper by the setup: its global index, its
// transformed vertices and the offset from those to its mesh vertices
// (colours, UVs).  Its fan pieces take the primitive slots
// n_tris + CLIP_PIECES * (queue position) + piece.
struct ClipJob {
    int32_t g, i0, i1, i2;
    int32_t dv, pad[3];
};
constexpr int CLIP_PIECES = 7;  // a triangle clipped by six planes keeps at most 9 vertices: 7 pieces

struct Counters {
    int32_t n_prims;
    int32_t n_clipped;      // triangles queued for the clipper
    int32_t n_pairs;        // (primitive, tile) pairs
    int32_t nan_flag;
    int32_t overflow;
    int32_t shadow_nan;
    int32_t n_chunks;       // raster work items: (tile, up to ITEM_PRIMS primitives)
    int32_t next_item;      // persistent raster work queue head
    int32_t depth_nonfinite;  // some pixel's resolved depth is NaN or infinite
    int32_t pad[7];
};

struct Capacities {
    int32_t prims;
    int32_t clip_subs;
    int32_t pairs;
    int32_t clip_tris;
};

constexpr int CHUNK = 256;          // primitives staged in LDS at once
constexpr int ITEM_PRIMS = 2048;    // primitives per raster work item (the default; prh_tuning.item_prims)

// Where a raster pass writes.
struct RasterTarget {
    int32_t W, H, tiles_x, tiles_per_map, n_tiles;
    unsigned long long* vis;  // main view: 64-bit depth|order keys
    uint32_t* first_all;      // exact-order pass: first fragment per pixel
    uint32_t* first_nan;      // exact-order pass: first NaN-depth fragment
    uint32_t* smap_ord;       // shadow maps as ordered depth bits
};

__host__ __device__ inline uint32_t pack_bbox(int x, int y) { return (uint32_t)(x & 0xffff) | ((uint32_t)y << 16); }

}  // namespace prh
