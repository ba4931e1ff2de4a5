/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of polyred's CPU renderer (the parity oracle for
 * libprh).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline; the
 * product path never calls it.
 *
 * It follows the reference in sequential Workers(1) order (one worker makes
 * internal/sched a single FIFO channel, internal/sched/sched.go:40-78), which
 * is the order the reference's own parity tests pin
 * (render/gpudeferred_compare_test.go:13-22).
 *
 * Pinning (each check lives in tests/test_oracle_*.py):
 *   - Barycoord, ViewportMatrix, ViewMatrix, Perspective/Orthographic
 *     ProjMatrix, AABB Intersect/Contains/Add, Triangle.IsValid,
 *     interpWorldPos, matAt, sRGB endpoints and LUT round trip, the shadow
 *     quantisation: against the reference's known-answer tests
 *     (math/interpolate_test.go:104-115, math/math_test.go:40-51,
 *     camera/camera_test.go:106-196, geometry/primitive/box_test.go:30-110,
 *     geometry/primitive/triangle_test.go:29-52, render/worldpos_test.go:17-31,
 *     render/material_table_test.go:17-34, color/color_test.go:36-67,
 *     gpu/shader/gpumath/kernels/shadow_test.go:12-31);
 *   - FragmentShader: the six fragments of
 *     render/shading_equiv_test.go:33-107 against an independent FP64
 *     Blinn-Phong, and a brute-force FP64 re-evaluation on whole frames;
 *   - the raster: brute force on tiny scenes (every pixel's coverage from an
 *     exact rational edge test), ties and clipping fans by construction;
 *   - shadows: the light-camera fit by the invariants it must satisfy
 *     (every main-camera frustum corner inside the light's NDC cube, the
 *     faces touched, near/far as the fit defines them); the maps texel by
 *     texel against exact-rational coverage of the float64 light camera;
 *     the lookup's rules (texel 0 skipped, truncation, row wrap, strict
 *     bias, transform order) on hand cases; an occluder's footprint, the
 *     zero-texel quirk and the non-casting-light count in closed form;
 *   - AO: every direction's largest elevation angle on step, ramp, border,
 *     truncation and uncovered-fragment depth buffers, and whole frames of
 *     a two-plane step, against closed forms of the geometry.  Go's
 *     math.Atan/Cos/Sin last bits are not reproducible without Go; the
 *     oracle rounds the C library's float64 results to float32.
 * Each check was also run against planted mistakes in this file (flipped
 * signs, floor for truncation, missing skips and culls, transposed
 * indices, reordered transforms), every one of which it rejects.
 */
#ifndef POLYRED_ORACLE_H
#define POLYRED_ORACLE_H

#include <stdint.h>
#include "../include/prh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_out {
    uint32_t* vis_order;   /* [H][W] raster rows: triangle*8+sub, 0xFFFFFFFF if not Ok */
    float* depth;          /* [H][W] fragment depth, -1 if not Ok                      */
    float* linear;         /* [H][W][3] pre-Round shading sums, NaN if not shaded      */
    uint8_t* pregamma;     /* [H][W][4] raster rows, colour after the deferred pass    */
    uint8_t* image;        /* [H][W][4] final image, top row first                     */
    float* shadow_maps;    /* [n_lights][S][S]                                          */
    const uint8_t* shade_mask; /* optional [H][W] raster rows: deferred-shade only these */
    double* shade_err64;   /* optional: max |f32 linear - f64 re-evaluation|            */
    uint8_t* shadow_count; /* optional [H][W] raster rows: the lights whose shadow test
                              reported true, at covered pixels whose material receives
                              shadows (the colour is scaled by 2^-count); 0 elsewhere  */
    float* ao_total;       /* optional [H][W] raster rows: AO's total before
                              pow(total, 10000) (material/ao.go:27-31), NaN where
                              no AO was applied                                       */
    float* light_matrices; /* optional [n_lights][2][16]: each shadow-casting
                              light's camera, view then projection (row-major);
                              untouched for the others                               */
} oracle_out;

/* With MSAA m > 1 every buffer is (height*m) x (width*m) except the final
 * image, which is resized to height x width (render/raster.go:361-378). */
/* Threads of the CPU baseline's timed frames: 0 (the default) renders in
 * the parity order (sequential raster; the deferred pass, whose result does
 * not depend on it, on OpenMP's threads); 1 runs everything on one thread
 * (Workers(1)); n > 1 draws the triangles of every raster pass on n threads
 * concurrently, each pixel's depth test and set under a spin lock as the
 * reference's FragmentBuffer does, and shades on n threads (ties between
 * concurrent triangles then resolve in arbitrary order, as in the reference
 * with several workers: never used for parity). */
void oracle_set_workers(int n);

int oracle_render(const prh_scene_desc* scene, const prh_camera* cam,
                  const prh_options* opts, oracle_out* out);

/* The FP64 reference of libprh's adjoint of the shading pass at fixed
 * visibility (oracle.c, "differentiable shading"); any gradient pointer
 * may be NULL, target NULL computes the forward only. */
typedef struct oracle_diff_params {
    const double* albedo;
    const double* normal;
    const double* kd;
    const double* ks;
    const double* shininess;
    const double* light_pos;
    const double* light_int;
    const uint8_t* shadow_count; /* optional [H][W]: the shading is scaled by 2^-count */
} oracle_diff_params;
typedef struct oracle_diff_grads {
    double *albedo, *normal, *kd, *ks, *shininess, *light_pos, *light_int;
} oracle_diff_grads;
int oracle_diff(const prh_scene_desc* s, const prh_camera* cam, const prh_options* o, const uint32_t* vis_order,
                const oracle_diff_params* p, const double* target, double* linear, oracle_diff_grads* g,
                double* loss);

/* imageutil.Resize (internal/imageutil/resize.go:16-66) of an iw x ih RGBA8
 * image to width x height. */
int oracle_resize(const uint8_t* img, int iw, int ih, uint8_t* out, int width, int height);

/* Texture.Query (buffer/texture.go:83-154) of a texture built as
 * buffer.NewTexture builds it (buffer/texture.go:35-68). */
int oracle_texture_query(const prh_texture* t, float lod, float u, float v, uint8_t out[4]);
/* Mip level `level` of that texture: its size into *w, *h and, when `out`
 * holds bytes >= w*h*4, its texels; returns the number of levels. */
int oracle_texture_level(const prh_texture* t, int level, int* w, int* h, uint8_t* out, size_t bytes);

/* Known-answer entry points (row-major Mat4 X00..X33). */
void oracle_barycoord(const float p[2], const float a[2], const float b[2],
                      const float c[2], float out[3]);
void oracle_viewport_matrix(float w, float h, float out[16]);
void oracle_view_matrix(const float pos[3], const float target[3],
                        const float up[3], float out[16]);
void oracle_proj_matrix(const prh_camera* cam, float out[16]);
void oracle_mat4_inv(const float m[16], float out[16]);
void oracle_mat4_mulm(const float a[16], const float b[16], float out[16]);
int oracle_aabb_intersect(const float min1[3], const float max1[3],
                          const float min2[3], const float max2[3]);
int oracle_aabb_contains(const float mn[3], const float mx[3], int n,
                         const float* pts);
void oracle_aabb_new(int n, const float* pts, float mn[3], float mx[3]);
int oracle_tri_is_valid(const float p1[4], const float p2[4], const float p3[4]);
void oracle_interp_worldpos(const float bc[3], const float m1[4],
                            const float m2[4], const float m3[4], float out[4]);
int oracle_mat_at(int32_t n, int64_t id);
float oracle_srgb_from_linear(float v);
float oracle_linear_from_srgb(float v);
uint8_t oracle_gamma_u8(uint8_t c);
uint8_t oracle_shadow_quant(uint8_t c, int occluded);
/* shader.FragmentShader on one fragment; lin may be NULL. */
void oracle_fragment_shader(const prh_material* m, const float nor[4],
                            const float facenor[4], const float wpos[4],
                            const float cam[3], int n_lights,
                            const prh_light* lights, int n_amb,
                            const float* amb, uint8_t out[4], float lin[3]);
/* AO direction table (8 x {cos, sin}) as material/ao.go:26 builds it. */
void oracle_ao_dirs(float out[16]);
/* maxElevationAngle for the eight directions and AmbientOcclusionShade's
 * total and factor (material/ao.go:20-73) at raster pixel (px, py) of a
 * W x H fragment buffer given as raster rows (index y*W + x): its depths and
 * Ok flags. */
int oracle_ao_march(const float* depth, const uint8_t* ok, int W, int H, int px, int py, float angles[8],
                    float* total, float* factor);
/* The lookup of shadingVisibility (render/shadow.go:243-282) of the fragment
 * (x, y, depth) in one SW x SH shadow map, with the given screen-to-world,
 * light view and light projection matrices: 1 if occluded. */
int oracle_shadow_lookup(const float s2w[16], const float lview[16], const float lproj[16], int SW, int SH,
                         const float* map, float x, float y, float depth);

#ifdef __cplusplus
}
#endif
#endif
