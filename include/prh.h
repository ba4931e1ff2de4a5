/*
 * prh.h — C ABI of libprh, the MI355X (gfx950) rasteriser and deferred shader
 * that sits behind polyred's render.NewRenderer(...).Render() option API.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int status (PRH_OK == 0, a negative PRH_E_* code otherwise).  Host memory
 * passed in is borrowed for the duration of the call only.  Handles are opaque.
 * No float is passed by value: purego's SyscallN carries integer and pointer
 * registers only, so every float travels inside a struct passed by pointer.
 *
 * Citations are to the polyred sources (paths relative to the polyred repo).
 * The Go-side binding a maintainer adds is shown in INTEGRATION.md.
 */
#ifndef PRH_H
#define PRH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PRH_OK 0
#define PRH_E_INVALID (-1)     /* bad argument or handle                        */
#define PRH_E_UNSUPPORTED (-2) /* scene feature not on the device path; the Go
                                  caller falls back to the CPU pass, as
                                  errGPUDeferredUnsupported does
                                  (render/gpudeferred.go:24, render/raster.go:67-76) */
#define PRH_E_HIP (-3)         /* a HIP runtime call failed                     */
#define PRH_E_NOMEM (-4)       /* device allocation failed                      */
#define PRH_E_NODEVICE (-5)    /* no such GPU                                   */

/* ---- scene contract (host-side, borrowed) ------------------------------ */

/* A texture image (buffer.NewTexture, buffer/texture.go:35-68): RGBA8 rows,
 * row 0 first as image.RGBA stores them.  With use_mipmap (TextureIsoMipmap,
 * true for textures a model loads) the library builds the mip levels as the
 * reference does, each an imageutil.Resize of the image. */
typedef struct prh_texture {
    int32_t width, height;
    const uint8_t* rgba;  /* [height][width][4] */
    int32_t use_mipmap;
} prh_texture;

/* Blinn-Phong material, one entry of the renderer's flat material table.
 * Mirrors material.BlinnPhong (material/material.go:44-52).  Its texture is
 * either textures[texture_id] of the scene or, with texture_id -1, the
 * uniform 1x1 texture of colour `texture` (buffer.NewUniformTexture,
 * buffer/texture.go:17-24), which Texture.Query returns for every
 * (lod, u, v) (buffer/texture.go:83-129). */
typedef struct prh_material {
    uint8_t diffuse[4];  /* BlinnPhong.Diffuse  RGBA8 */
    uint8_t specular[4]; /* BlinnPhong.Specular RGBA8 */
    uint8_t texture[4];  /* uniform texture colour RGBA8 */
    float shininess;     /* BlinnPhong.Shininess */
    int32_t flat_shading;      /* Standard.FlatShading     */
    int32_t ambient_occlusion; /* Standard.AmbientOcclusion */
    int32_t receive_shadow;    /* Standard.ReceiveShadow   */
    int32_t texture_id;        /* index into prh_scene_desc.textures, or -1 */
} prh_material;

#define PRH_LIGHT_POINT 0
#define PRH_LIGHT_DIRECTIONAL 1

/* A light.Source in Scene.Lights() order (scene/scene.go:35-47).
 * Point: light/point.go:22-46.  Directional: light/directional.go:23-48; the
 * direction is passed already normalised, as NewDirectional stores it
 * (light/directional.go:45).  A shadow-casting directional light uses the same
 * orthographic light camera as a point light, placed at `position` and looking
 * at the scene centre (render/shadow.go:33-90; the reference builds no camera
 * for directional lights, see the design notes). */
typedef struct prh_light {
    int32_t kind;        /* PRH_LIGHT_POINT or PRH_LIGHT_DIRECTIONAL */
    float position[3];
    float direction[3];
    float intensity;
    uint8_t color[4];
    int32_t cast_shadow;
} prh_light;

/* Scene geometry in the order render.cpuForwardPass visits it
 * (render/raster.go:226-273): objects in scene.IterObjects order
 * (scene/core.go:86-111), each object's triangles in Geometry.Triangles()
 * order.  Meshes hold vertices and triangles; an object instances one mesh
 * with its composed model matrix (root . groups . geometry, as
 * `modelMatrix.MulM(g.ModelMatrix())` at render/raster.go:242) and the base of
 * its materials in the flat table (render/raster.go:252-262). */
typedef struct prh_scene_desc {
    int32_t n_meshes;
    const int32_t* mesh_vtx_off; /* [n_meshes+1] into the vertex arrays        */
    const int32_t* mesh_tri_off; /* [n_meshes+1] into the triangle arrays      */
    const float* vpos;           /* [V][4] model-space position (x,y,z,w)     */
    const float* vnor;           /* [V][4] model-space normal (x,y,z,w)       */
    const uint8_t* vcol;         /* [V][4] vertex colour RGBA8                */
    const float* vuv;            /* [V][2] texture coordinates                */
    const int32_t* tri_v;        /* [T][3] mesh-local vertex indices          */
    const int32_t* tri_mat;      /* [T] geometry-local material id, <0 = vertex colour */

    int32_t n_objects;
    const int32_t* obj_mesh;     /* [O] mesh index                            */
    const float* obj_model;      /* [O][16] row-major Mat4 X00..X33           */
    const int32_t* obj_mat_base; /* [O] base of the object's materials        */

    int32_t n_materials;
    const prh_material* materials; /* the flat material table               */

    int32_t n_lights;
    const prh_light* lights;       /* light sources in Scene.Lights() order */
    int32_t n_ambient;
    const float* ambient_intensity; /* light.Ambient intensities in order   */

    float center[3];               /* Scene.Center() (scene/scene.go:49-52) */

    int32_t n_textures;
    const prh_texture* textures;   /* images the materials' texture_id name */
} prh_scene_desc;

#define PRH_CAMERA_PERSPECTIVE 0
#define PRH_CAMERA_ORTHOGRAPHIC 1

/* camera.Perspective (camera/perspective.go:22-111) or camera.Orthographic
 * (camera/orthographic.go:22-117). */
typedef struct prh_camera {
    int32_t kind;
    float position[3];
    float target[3];
    float up[3];
    float fov, aspect;                /* perspective: degrees, width/height */
    float near_, far_;                /* both kinds                         */
    float left, right, bottom, top;   /* orthographic                       */
} prh_camera;

/* render.option fields the device path honours (render/options.go:16-33). */
typedef struct prh_options {
    int32_t width, height;   /* render.Size                                    */
    int32_t msaa;            /* render.MSAA, 1..16: every pass at (w*m, h*m),
                                then imageutil.Resize                          */
    uint8_t background[4];   /* render.Background                              */
    int32_t shadow_map;      /* render.ShadowMap                               */
    int32_t gamma_correct;   /* render.GammaCorrection                         */
    int32_t shadow_size;     /* shadow map side; 0 = framebuffer-sized as the
                                reference (render/shadow.go:87-88)             */
    int32_t keep_linear;     /* also keep the debug buffers: pre-Round linear
                                shading sums and the pre-gamma colour          */
} prh_options;

/* ---- handles ----------------------------------------------------------- */
typedef struct prh_ctx prh_ctx;
typedef struct prh_scene prh_scene;
typedef struct prh_renderer prh_renderer;

/* Threads: every call may come from any OS thread (each sets its context's
 * device first, as goroutines migrate).  Calls on one context -- its
 * scenes, renderers, fits and surfaces -- must not overlap (the Go side
 * renders from one goroutine); different contexts may be used from
 * different threads at once. */

/* Thread-local text of the last error on the calling thread. */
const char* prh_last_error(void);

/* Open GPU `ordinal` and create the context's stream.  Replaces gpu.Open
 * (gpu/device.go:77-89) for this option; the caller owns the context.
 * Lifetimes nest: a context outlives its scenes and a scene its renderers;
 * prh_close / prh_scene_free return PRH_E_INVALID while children live. */
int prh_open(int32_t ordinal, prh_ctx** out);
int prh_close(prh_ctx* ctx);

/* Copy a scene into device memory (HBM).  The descriptor is borrowed.
 * Replaces the per-frame CPU marshalling of render/gpuforward.go:271-310 and
 * render/gpudeferred.go:130-187 with one upload. */
int prh_scene_upload(prh_ctx* ctx, const prh_scene_desc* desc, prh_scene** out);
int prh_scene_free(prh_scene* scene);

/* Create a renderer: render.NewRenderer(Size, Camera, Scene, Background,
 * ShadowMap, GammaCorrection, ...) (render/raster.go:84-143), including the
 * one-time shadow-camera fit of initShadowMaps (render/shadow.go:33-90) and
 * its zero-initialised shadow maps. */
int prh_renderer_new(prh_ctx* ctx, prh_scene* scene, const prh_camera* camera,
                     const prh_options* opts, prh_renderer** out);
int prh_renderer_free(prh_renderer* r);

/* One frame of (*Renderer).Render() (render/raster.go:155-199): shadow
 * passes, forward raster, deferred shading, gamma.  Returns when the frame is
 * complete on the device; the image stays in HBM until read. */
int prh_render(prh_renderer* r);

/* One frame of each of n renderers (all of one context), as n prh_render
 * calls would produce them -- shadow passes included -- with no host
 * synchronisation between views: by default one launch per pass covers
 * every view (each view's workgroups find its arguments in a table built
 * once per batch), with PRH_BATCH_STREAMS each view's own launches over the
 * context's batch streams.  Each view's counters are checked once at the
 * end and a view that overflowed its work buffers is rendered again on its
 * own.  With PRH_BATCH_GRAPH the whole batch is captured once into a HIP
 * graph (kept until the batch changes) and replayed.  Per-pass statistics
 * are not recorded.  *ms (may be NULL) receives the batch's device time.  No
 * counterpart in polyred (one renderer draws one view). */
#define PRH_BATCH_GRAPH 1
#define PRH_BATCH_STREAMS 2
int prh_render_batch(prh_renderer* const* renderers, int32_t n, int32_t flags, float* ms);

/* The rendered RGBA8 image, top row first, the layout of
 * FragmentBuffer.Image() (buffer/buffer.go:160-166).  bytes >= 4*W*H. */
int prh_read_image(prh_renderer* r, uint8_t* rgba, size_t bytes);

#define PRH_BUF_VIS 0       /* uint64 [H][W] raster rows: depth|primitive key  */
#define PRH_BUF_DEPTH 1     /* float  [H][W] fragment depth, -1 if uncovered   */
#define PRH_BUF_LINEAR 2    /* float  [H][W][3] pre-Round shading sums         */
#define PRH_BUF_SHADOW 3    /* float  [n_lights][S][S] shadow maps             */
#define PRH_BUF_PREGAMMA 4  /* uint8  [H][W][4] deferred colour before gamma   */
#define PRH_BUF_SHADOW_COUNT 5 /* uint8 [H][W] lights whose shadow test reported
                                  true at covered pixels whose material receives
                                  shadows (the colour is scaled by 2^-count)   */

/* Copy one internal buffer to host memory (parity tests, debugging). */
int prh_read_buffer(prh_renderer* r, int32_t which, void* dst, size_t bytes);

/* Per-kernel device time of the last frame, from HIP events on the
 * renderer's stream. */
typedef struct prh_frame_stats {
    float ms_total;
    float ms_shadow;     /* all shadow passes                               */
    float ms_vertex;     /* vertex transform of the main view               */
    float ms_setup;      /* cull, clip, bin                                 */
    float ms_raster;     /* visibility raster                               */
    float ms_shade;      /* deferred shade + shadow lookups + gamma (k_shade) */
    float ms_ao;         /* screen-space ambient occlusion                  */
    int32_t n_prims;     /* primitives after cull/clip: unclipped triangles
                            plus the clipper's pieces                       */
    int32_t n_clipped;   /* triangles that went through the clipper         */
    int32_t n_bin_pairs; /* (primitive, tile) pairs                         */
    int32_t nan_rerun;   /* 1 if the NaN-depth exact-order pass ran         */
    float ms_diff;       /* the last prh_diff_forward / _backward kernel    */
} prh_frame_stats;
int prh_last_stats(prh_renderer* r, prh_frame_stats* out);

/* ---- differentiable shading (no counterpart in polyred) ---------------
 * The forward and the adjoint of the deferred Blinn-Phong shading at the
 * fixed visibility of the renderer's last frame, for gradient-based inverse
 * rendering (the Go side exposes it as Renderer.Backward).  The shading is
 * FragmentShader's sums before Round (shader/blinn_cpu.go:24-106) with
 * real-valued parameters: a per-vertex albedo interpolated perspective-
 * correctly (in place of the uniform texture colour), per-vertex normals,
 * per-material Kd, Ks (0..255 scale) and shininess, per-light position and
 * intensity.  The loss is the sum over shaded pixels and RGB of
 * (linear - target)^2.  Buffers are MSAA-sized raster rows.  At most 8
 * lights (PRH_E_UNSUPPORTED above). */
typedef struct prh_diff_params {
    const float* albedo;    /* [n_vertices][3] per transformed vertex            */
    const float* normal;    /* [n_vertices][3], or NULL: the scene's normals
                               transformed by each object's normal matrix       */
    const float* kd;        /* [n_materials][3]                                  */
    const float* ks;        /* [n_materials][3]                                  */
    const float* shininess; /* [n_materials]                                     */
    const float* light_pos; /* [n_lights][3] (directional lights: unused)        */
    const float* light_int; /* [n_lights]                                        */
} prh_diff_params;
typedef struct prh_diff_grads { /* host outputs with prh_diff_params' shapes; any may be NULL */
    float *albedo, *normal, *kd, *ks, *shininess, *light_pos, *light_int;
} prh_diff_grads;

/* Transformed vertices (objects x mesh vertices, object order), materials
 * and lights of the renderer's scene. */
int prh_diff_sizes(prh_renderer* r, int32_t* n_vertices, int32_t* n_materials, int32_t* n_lights);
/* Linear shading of every pixel: float [H][W][3], NaN where not shaded. */
int prh_diff_forward(prh_renderer* r, const prh_diff_params* p, float* linear, size_t bytes);
/* Loss and its gradients; target is float [H][W][3]. */
int prh_diff_backward(prh_renderer* r, const prh_diff_params* p, const float* target, size_t bytes, double* loss,
                      const prh_diff_grads* grads);
/* prh_diff_backward without any device allocation.  No counterpart in
 * polyred: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or optimisation loops, hosts that forbid
 * hipMalloc).  Same arguments, kernels and results as prh_diff_backward, but
 * the parameters, gradients, target, loss and shadow counts live in one
 * static device array of PRH_NOALLOC_SCRATCH_BYTES per GPU (held from the
 * library's load until the process ends) instead of the renderer's own
 * buffers; a frame needing more returns PRH_E_INVALID (about 12 bytes per
 * pixel plus 48 per transformed vertex).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (from several threads or contexts)
 * race, so serialise them.  The renderer's frame buffers are not touched, and
 * prh_last_stats' ms_diff is not updated. */
#define PRH_NOALLOC_SCRATCH_BYTES (32u << 20)
int prh_diff_backward_noalloc(prh_renderer* r, const prh_diff_params* p, const float* target, size_t bytes,
                              double* loss, const prh_diff_grads* grads);

/* ---- batched inverse rendering (no counterpart in polyred) -------------
 * Gradient descent on the differentiable shading above over many views of
 * one scene at once.  Every view's last frame fixes its visibility and its
 * shadow visibility: the shading of a pixel is scaled by 2^-count, count the
 * lights whose shadow test (render/shadow.go:230-283) reported true there,
 * as the reference scales the shaded colour (render/raster.go:339-347).  One
 * step is one launch of the adjoint over every view (gradients summed over
 * views) and one Adam update (Kingma & Ba 2015) of the chosen parameter
 * groups, parametrised as p = start * (1 + theta), theta starting at 0; the
 * steps of a run are one captured HIP graph replayed.  The loss is the sum
 * over views of prh_diff_backward's loss.  kd, ks, shininess, light
 * intensities and albedo are kept >= 0. */
typedef struct prh_fit prh_fit;
#define PRH_FIT_ALBEDO 1 /* (normals are not fitted) */
#define PRH_FIT_KD 4
#define PRH_FIT_KS 8
#define PRH_FIT_SHININESS 16
#define PRH_FIT_LIGHT_POS 32
#define PRH_FIT_LIGHT_INT 64
/* prh_fit_new flag: read the adjoint's per-pixel state from a G-buffer
 * stored once per run (64 bytes a pixel) instead of rebuilding it from the
 * visibility buffer at every step; albedo gradients need the rebuild
 * (PRH_E_UNSUPPORTED).  Start and truth parameters take no normals. */
#define PRH_FIT_STORED_GBUFFER 256
typedef struct prh_fit_options {
    int32_t steps;
    int32_t groups; /* PRH_FIT_ALBEDO .. PRH_FIT_LIGHT_INT to optimise */
    float lr, beta1, beta2, eps;
} prh_fit_options;
typedef struct prh_fit_stats {
    double visibility_bytes; /* the fixed visibility an adjoint step reads: the
                                visibility buffer and shadow counts, or the
                                stored G-buffer, summed over views          */
    float ms_steps;          /* the run's steps (HIP events)                  */
    float ms_gbuffer;        /* building the stored G-buffer                  */
    int32_t steps;
    int32_t n_views;
} prh_fit_stats;
/* Views of one scene, each with a rendered frame (prh_render or
 * prh_render_batch), each at most once; rendering a view again invalidates
 * the fit, and a view cannot be freed before the fits that hold it. */
int prh_fit_new(prh_renderer* const* views, int32_t n, int32_t flags, prh_fit** out);
int prh_fit_free(prh_fit* f);
/* Every view's target := its forward shading (prh_diff_forward, shadow
 * scale included) at `truth`, on the device. */
int prh_fit_targets(prh_fit* f, const prh_diff_params* truth);
/* View `view`'s target from the host: float [H][W][3]. */
int prh_fit_set_target(prh_fit* f, int32_t view, const float* target, size_t bytes);
/* o->steps steps from `start`; `result` (arrays of prh_diff_params'
 * shapes, any may be NULL) receives the final parameters and
 * loss_history[i] (may be NULL) the loss before step i. */
int prh_fit_run(prh_fit* f, const prh_diff_params* start, const prh_fit_options* o, const prh_diff_grads* result,
                double* loss_history);
int prh_fit_last_stats(prh_fit* f, prh_fit_stats* out);

/* ---- headless present (gpu/surface.go:1-183) --------------------------
 * gpu.Surface's headless swapchain: a ring of `frames` RGBA8 textures in HBM
 * (top row first, as prh_read_image), one per frame, that a renderer draws
 * into; the seam an on-screen present builds on. */
typedef struct prh_surface prh_surface;
/* Device.CreateSurface (surface.go:43-57): frames <= 0 means 2. */
int prh_surface_new(prh_ctx* ctx, int32_t width, int32_t height, int32_t frames, prh_surface** out);
int prh_surface_free(prh_surface* s); /* Release */
/* AcquireNextTexture (surface.go:102-110): *index (may be NULL) receives the
 * ring slot the next frame is drawn into. */
int prh_surface_acquire(prh_surface* s, int32_t* index);
/* One prh_render of r whose final image is written into the acquired
 * texture instead of r's own image (r's output size must be the surface's). */
int prh_surface_render(prh_surface* s, prh_renderer* r);
/* PresentImage's upload (surface.go:133-137): rgba into the acquired texture. */
int prh_surface_write(prh_surface* s, const uint8_t* rgba, size_t bytes);
/* Present (surface.go:115-129): completes the acquired frame; an error
 * without an acquire. */
int prh_surface_present(prh_surface* s);
/* Texture().ReadPixels(): the most recently presented frame's pixels. */
int prh_surface_read(prh_surface* s, uint8_t* rgba, size_t bytes);
/* Resize (surface.go:163-177): new textures, the ring restarts. */
int prh_surface_resize(prh_surface* s, int32_t width, int32_t height);
int prh_surface_size(prh_surface* s, int32_t* width, int32_t* height);

/* ---- autotuning (no counterpart in polyred) ------------------------------
 * Knobs that change how fast a frame renders but not what it renders: every
 * merge they touch is order-independent, so frames are bit-identical under
 * any setting.  The tile is fixed at 16x16: one 256-thread workgroup holds a
 * tile with one pixel per thread, and the shading's 8x8 wave blocks and the
 * AO's row segments are laid out on it. */
typedef struct prh_tuning {
    int32_t raster_wg_per_cu;  /* persistent raster workgroups per CU (occupancy), 1..16; default 8        */
    int32_t item_prims;        /* primitives per raster work item (bin capacity), a power of 2 in
                                  64..8192; default 2048                                                   */
    int32_t shadow_direct_max; /* shadow-map triangles of at most this many texel centres are drawn by
                                  the setup, 0..256; default 64                                            */
    int32_t main_direct_max;   /* main-view triangles of at most this many pixel centres, likewise; 48  */
    int32_t overlap_shadows;   /* 1: a frame's shadow passes run on a second stream beside the main
                                  view's passes; default 1 from 2^19 framebuffer pixels, else 0          */
    int32_t raster_tile;       /* side of the binning and raster tiles (main view and shadow maps):
                                  8, 16 or 32 pixels; default 16                                          */
    float ms;                  /* prh_autotune: the chosen configuration's median frame time             */
} prh_tuning;
int prh_get_tuning(prh_renderer* r, prh_tuning* out);
int prh_set_tuning(prh_renderer* r, const prh_tuning* t);
/* Coordinate descent over item_prims {256..4096}, raster_wg_per_cu {1, 2, 4,
 * 8}, shadow_direct_max {0, 16, 32, 64, 128, 256}, main_direct_max {0, 24,
 * 48, 96, 192}, overlap_shadows {0, 1} and raster_tile {8, 16, 32}: each
 * configuration timed as the median of
 * `frames` frames (HIP events, after one warm-up frame); the winner is timed
 * again against the starting knobs over twice the frames and the faster of
 * the two stays on r and goes to *out (may be NULL).  r holds a frame of it
 * on return. */
int prh_autotune(prh_renderer* r, int32_t frames, prh_tuning* out);
/* The same search for a batch (no counterpart in polyred): the views of a
 * prh_render_batch of wide launches, split into two size classes (class 0
 * below 2^19 framebuffer pixels, class 1 the rest), each class with its own
 * item_prims, raster_wg_per_cu (in a wide batch: n_CU * raster_wg_per_cu /
 * 16 persistent raster workgroups per view, at least 32), shadow_direct_max,
 * main_direct_max and raster_tile.  Coordinate descent over the (class, knob) pairs, each
 * configuration timed as the median of `frames` eager wide batches; the
 * winner is checked against the starting knobs as prh_autotune does.  Every
 * view keeps its class's winner; out[0..1] (may be NULL) receive the two,
 * with ms the batch's median time. */
int prh_autotune_batch(prh_renderer* const* renderers, int32_t n, int32_t frames, prh_tuning out[2]);

/* Page-locked host memory for the read calls (prh_read_image,
 * prh_read_buffer, prh_surface_read): a copy into it runs at the link's
 * full rate, where a copy into pageable memory is staged by the runtime. */
int prh_host_alloc(size_t bytes, void** out);
int prh_host_free(void* p);

/* Library build identification ("libprh gfx950 ..."). */
const char* prh_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PRH_H */
