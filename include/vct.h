/*
 * vct.h - C ABI of the MI355X (gfx950) voxel cone tracing core.
 *
 * This is the drop-in boundary that replaces the three OpenGL passes of
 * latencyhiding/voxel_cone_tracing (voxelize, anisotropic mip, cone trace).
 * Every entry point names the reference interface it stands in for
 * (reference file:line).  The C++ `Renderer` / `create_tex_3d` shim in
 * include/vct_renderer.h keeps the reference's names on top of this ABI.
 *
 * Conventions
 *  - Every call returns int: VCT_OK (0) or a negative VCT_ERR_* code; the
 *    message of the last failure on the calling thread is vct_last_error().
 *  - Host arrays passed in are copied, never retained.
 *  - Matrices are 4x4 float, column-major (the memory layout of glm::mat4
 *    that the reference passes to glUniformMatrix4fv, renderer.cpp:244).
 *  - One context per host thread; all device work of a context runs in
 *    order on the context's own HIP stream.  Calls return once the work is
 *    queued; vct_sync() waits for it.
 *  - No GPU fallback: a context cannot be created without a gfx950 device.
 */
#ifndef VCT_H
#define VCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define VCT_OK                 0
#define VCT_ERR_INVALID_ARG   -1
#define VCT_ERR_BAD_HANDLE    -2
#define VCT_ERR_OUT_OF_MEMORY -3
#define VCT_ERR_HIP           -4
#define VCT_ERR_NO_DEVICE     -5
#define VCT_ERR_STATE         -6
#define VCT_ERR_LIMIT         -7

#define VCT_INVALID_ID 0xFFFFFFFFu   /* INVALID_ID (~0), renderer.h:22 */
#define VCT_MAX_LIGHTS 10            /* MAX_POINT_LIGHTS, voxel_cone_tracing.frag:52 */

/* Voxel storage type of the radiance volume and its mips.
 * VCT_RGBA8 is the reference's GL_RGBA8 (texture_3d.cpp:17); the others
 * store linear radiance in 16/32-bit floats, accumulated in FP32. */
typedef enum {
  VCT_RGBA8 = 0,
  VCT_FP16 = 1,
  VCT_BF16 = 2,
  VCT_FP32 = 3
} vct_dtype;

/* How colliding fragments are combined into one level-0 voxel.
 * VCT_AVG_REFERENCE: the packed RGBA8 running average with the 4-bit count
 *   nibble of imageAtomicRGBA8Avg (voxelize.frag:95-120), replayed per voxel
 *   in one canonical fragment order (triangle id, then pixel quad, then
 *   pixel), which makes it deterministic.  RGBA8 storage only.
 * VCT_AVG_EXACT: exact mean of the fragment colours (16.16 fixed point
 *   sums, order independent).  Any storage type. */
typedef enum {
  VCT_AVG_REFERENCE = 0,
  VCT_AVG_EXACT = 1
} vct_average;

/* Diffuse cone sets.  VCT_CONES_REF9 is trace_diffuse
 * (voxel_cone_tracing.frag:140-168): the normal plus 8 cones mixed half way
 * to the tangent frame, aperture 0.55785173935. */
typedef enum {
  VCT_CONES_REF9 = 0,
  VCT_CONES_5_WIDE = 1,  /* normal + 4 side cones, aperture 2*tan(30 deg) */
  VCT_CONES_16 = 2       /* normal + rings of 5 and 10, aperture 0.45 */
} vct_cone_set;

typedef enum {
  VCT_LIGHT_POINT = 0,       /* point_light_t, renderer.h:87-92 */
  VCT_LIGHT_DIRECTIONAL = 1  /* position[] holds the direction towards the light */
} vct_light_kind;

/* material_data_t (renderer.h:94-120): packed, std140-compatible, 128 B. */
typedef struct vct_material {
  float ambient[4];
  float diffuse[4];
  float specular[4];
  float transmittance[4];
  float emission[3];
  float shininess;
  float ior;
  float dissolve;
  int32_t illum;
  float roughness;
  float metallic;
  float sheen;
  float clearcoat_thickness;
  float clearcoat_roughness;
  float anisotropy;
  float anisotropy_rotation;
  float pad[2];
} vct_material;

/* point_light_t (renderer.h:87-92) plus a kind tag. */
typedef struct vct_light {
  float position[3];
  float color[3];
  float intensity;
  int32_t kind;
} vct_light;

typedef struct vct_context* vct_ctx;

/* ---- context (Renderer::Renderer / ~Renderer, renderer.h:124-126) ---- */
const char* vct_last_error(void);
int vct_device_count(int* out);
/* Renderer(int width, int height), renderer.cpp:100-147: viewport w x h,
 * grid 128^3 with 7 levels (renderer.cpp:146,186), RGBA8 storage,
 * reference averaging, all rendering phases on, view_voxel_dir 0. */
int vct_create(int width, int height, vct_ctx* out);
int vct_destroy(vct_ctx ctx);
int vct_sync(vct_ctx ctx);
int vct_set_viewport(vct_ctx ctx, int width, int height);

/* ---- grid configuration ---------------------------------------------- */
/* Renderer::set_grid_resolution (renderer.cpp:178-188): res^3 voxels,
 * 7 levels (the reference's fixed count). res must be a power of two in
 * [4, 1024]. */
int vct_set_grid_resolution(vct_ctx ctx, unsigned res);
/* Extension of set_grid_resolution: levels in [1, log2(res)+1], storage
 * dtype and averaging mode. */
int vct_set_grid_format(vct_ctx ctx, unsigned res, int levels, int dtype, int average);
/* Renderer::set_grid_size (renderer.cpp:190-193): world half-extent. */
int vct_set_grid_size(vct_ctx ctx, float size);
int vct_get_grid(vct_ctx ctx, int* res, int* levels, int* dtype, int* average);

/* ---- camera (Renderer::set_camera_transform, renderer.cpp:167-171) ---- */
int vct_set_camera_transform(vct_ctx ctx, const float view[16], const float projection[16]);

/* ---- materials (renderer.cpp:643-665) -------------------------------- */
/* add_material: material 0 is the default material (renderer.cpp:115-128). */
int vct_add_material(vct_ctx ctx, const vct_material* m, uint32_t* out_id);
/* Bulk add of n materials with consecutive ids (per-triangle albedo). */
int vct_add_materials(vct_ctx ctx, const vct_material* m, uint32_t n, uint32_t* first_id);
int vct_upload_material_data(vct_ctx ctx, const vct_material* m, uint32_t id);
int vct_get_material(vct_ctx ctx, uint32_t id, vct_material* out);

/* ---- models (load_model upload part, renderer.cpp:559-602) ----------- */
/* Non-indexed triangle list: positions and per-vertex normals are 9 floats
 * per triangle (v0,v1,v2 xyz).  tri_material (n entries or NULL = all 0)
 * holds material ids; a model is one draw object per material run. */
int vct_model_create(vct_ctx ctx, const float* positions, const float* normals,
                     const uint32_t* tri_material, uint32_t ntri, uint32_t* out_id);
/* load_model (renderer.cpp:407-603) for Wavefront OBJ/MTL files:
 * triangulated faces, per-face materials from the MTL.  Returns
 * VCT_INVALID_ID in *out_id on a read failure, like the reference. */
int vct_load_model(vct_ctx ctx, const char* filename, uint32_t* out_id);
/* The parsing half of load_model (renderer.cpp:407-537) without a context
 * or a GPU: vct_obj_open reads the OBJ and the MTL files it names and
 * reports the triangle and material counts; vct_obj_read copies positions
 * and normals (9 floats per triangle), each triangle's index into the file's
 * materials (-1: none), the materials, and the model dimensions as the
 * reference computes them (renderer.cpp:457-486, 517). */
typedef struct vct_obj_data* vct_obj;
#define VCT_ERR_NOT_FOUND -8
int vct_obj_open(const char* filename, vct_obj* out, uint32_t* ntri, uint32_t* nmat);
int vct_obj_read(vct_obj obj, float* positions, float* normals, int32_t* tri_material, vct_material* materials,
                 float dims[3]);
void vct_obj_close(vct_obj obj);
int vct_set_model_material(vct_ctx ctx, uint32_t material_id, uint32_t model_id);
int vct_set_model_transform(vct_ctx ctx, uint32_t model_id, const float m[16]);
int vct_get_model_dimensions(vct_ctx ctx, uint32_t model_id, float out[3]);
int vct_model_triangle_count(vct_ctx ctx, uint32_t model_id, uint32_t* out);

/* ---- per-frame queues (renderer.cpp:209-223) -------------------------- */
int vct_queue_model(vct_ctx ctx, uint32_t model_id);          /* bad ids ignored */
int vct_queue_point_light(vct_ctx ctx, const vct_light* light); /* >10 ignored by the shaders */
int vct_clear_queues(vct_ctx ctx);

/* ---- render settings (renderer.cpp:195-207) -------------------------- */
int vct_set_rendering_phases(vct_ctx ctx, int direct, int diffuse, int specular, int shadow);
int vct_set_voxel_view_dir(vct_ctx ctx, int dir, float lod);
int vct_set_cone_set(vct_ctx ctx, int cone_set);

/* ---- passes ------------------------------------------------------------ */
/* Renderer::voxelize (renderer.cpp:316-348 without the filter() call):
 * clears level 0 and scatters every queued model into it. */
int vct_voxelize(vct_ctx ctx);
/* Renderer::filter (renderer.cpp:283-314): levels 1..L-1 of the six
 * direction maps from level 0 (mipmap.comp:45-99). */
int vct_mip_aniso(vct_ctx ctx);
/* The rasterisation half of Renderer::visualize (renderer.cpp:355-390,
 * voxel_cone_tracing.vert:22-28): per-pixel world position, normal,
 * linear depth and material of the nearest surface. */
int vct_gbuffer(vct_ctx ctx);
/* The shading half of Renderer::visualize: voxel_cone_tracing.frag:246-275
 * for every covered pixel; background keeps the clear colour
 * (0.15, 0.25, 0.25, 1) of renderer.cpp:398. */
int vct_trace(vct_ctx ctx);
/* Renderer::render (renderer.cpp:392-405): voxelize, mip, gbuffer, trace,
 * then the queues are cleared. */
int vct_render(vct_ctx ctx);

/* ---- read back / inject (tests, tools) ---------------------------------- */
/* Level `level` (0..L-1) of direction map `dir` (0..5: -x,+x,-y,+y,-z,+z),
 * in x-fastest linear order, elements of the storage dtype (RGBA8: uint32
 * with R in bits 0-7; FP16/BF16: 4 x uint16; FP32: 4 x float).
 * Level 0 is shared by the six maps (dir is ignored there). */
size_t vct_volume_level_bytes(vct_ctx ctx, int level);
int vct_volume_download(vct_ctx ctx, int level, int dir, void* host, size_t bytes);
int vct_volume_upload(vct_ctx ctx, int level, int dir, const void* host, size_t bytes);
/* Per-pixel G-buffer, row 0 = bottom row (GL window convention):
 * pos_depth = world xyz + view-space distance along -z (0 = empty pixel);
 * normal_mat = interpolated normal xyz + material id as float bits. */
int vct_gbuffer_download(vct_ctx ctx, float* pos_depth, float* normal_mat, size_t pixels);
int vct_gbuffer_upload(vct_ctx ctx, const float* pos_depth, const float* normal_mat, size_t pixels);
/* RGBA32F image of the last trace, row 0 = bottom row. */
int vct_image_download(vct_ctx ctx, float* rgba, size_t pixels);
/* Milliseconds of the last frame's passes measured with HIP events on the
 * context stream: [voxelize, mip, gbuffer, trace, upsample, fallback]
 * (n <= 6 entries are written); 0 for a pass that ran untimed. */
int vct_get_pass_times(vct_ctx ctx, float* ms, int n);
/* Which passes record timing events: bit k for pass k of vct_get_pass_times
 * (default 0x3f, all).  Each event is a packet the stream processes between
 * kernels (~3-4 us of idle GPU each on the 1M-triangle frame: 28 us of the
 * plain frame, 40 us of the learned one); 0 turns the diagnostic off.  No
 * counterpart in the reference, which does not time its passes. */
int vct_set_pass_timing(vct_ctx ctx, uint32_t mask);
/* Fragments and candidate pixels of the last voxelize. */
int vct_get_voxelize_stats(vct_ctx ctx, uint64_t* fragments, uint64_t* candidates);
/* Runs vct_trace once more with a counter and returns the number of cone
 * steps taken (each step = 3 textureLod samples of sample_voxel,
 * voxel_cone_tracing.frag:80-86): the unit of the effective sample
 * bandwidth.  Diagnostics; not part of a frame. */
int vct_trace_count_steps(vct_ctx ctx, uint64_t* steps);
/* Diagnostics with no counterpart in the reference: the peak rate (GB/s) at
 * which this GPU's vector memory path delivers 16-byte-per-lane loads that hit
 * the L1 cache, measured with a calibration kernel in the marcher's load form.
 * It is the ceiling of the marcher's texel gathers (their roofline). */
int vct_trace_load_peak(vct_ctx ctx, double* gbps);

/* ---- Texture3D (texture_3d.h:6-10) --------------------------------------- */
/* create_tex_3d (texture_3d.cpp:3-25): levels of a w x h x d texture,
 * CLAMP_TO_BORDER, LINEAR_MIPMAP_LINEAR; storage dtype selectable. */
int vct_tex3d_create(vct_ctx ctx, int w, int h, int d, int levels, int dtype, uint32_t* out);
int vct_tex3d_destroy(vct_ctx ctx, uint32_t tex);                 /* texture_3d.cpp:27-30 */
int vct_tex3d_activate(vct_ctx ctx, uint32_t tex, uint32_t unit); /* texture_3d.cpp:32-37 (binding only) */
int vct_tex3d_clear(vct_ctx ctx, uint32_t tex, const float rgba[4]); /* texture_3d.cpp:39-44, level 0 */
int vct_tex3d_mip(vct_ctx ctx, uint32_t tex);                     /* texture_3d.cpp:46-51, box filter */
int vct_tex3d_upload(vct_ctx ctx, uint32_t tex, int level, const void* host, size_t bytes);
int vct_tex3d_download(vct_ctx ctx, uint32_t tex, int level, void* host, size_t bytes);

/* ---- scene generators (host side, deterministic) ----------------------- */
/* Random triangle soup: centres U[-extent, extent]^3, equilateral with edge
 * LogNormal(mu, sigma), uniformly random orientation, per-vertex normal =
 * face normal.  Writes 9 floats per triangle to pos and nrm, and 3 floats
 * (albedo, U[0,1]^3) plus 1 float (roughness, U[rmin, rmax]) per triangle
 * to albedo_rough (4 floats per triangle). */
int vct_scene_soup(uint64_t seed, uint32_t ntri, float extent, float mu, float sigma,
                   float rmin, float rmax, float* pos, float* nrm, float* albedo_rough);
/* Procedural Cornell box: 5 quads of the unit room [-0.5,0.5]^3 open at +z
 * (red -x wall, green +x wall, white floor/ceiling/back) and two boxes
 * rotated about y; 34 triangles.  tri_material gets ids 0..4 relative to
 * the 5 materials written to mats (albedo of the boxes from the seed). */
int vct_scene_cornell(uint64_t seed, float* pos, float* nrm, uint32_t* tri_material,
                      vct_material* mats);

/* ---- learned quarter-resolution upsampler --------------------------------- */
/* A per-pixel MLP 34 -> 64 -> 64 -> 3 (ReLU), BF16 weights, FP32
 * accumulation.  Features of a full-resolution pixel: the 3x3 neighbourhood
 * of the quarter-resolution indirect irradiance around it (27), its normal
 * (3), linear depth (1) and albedo (3).  It has no counterpart in the
 * reference. */
#define VCT_MLP_IN 34
#define VCT_MLP_HIDDEN 64
#define VCT_MLP_OUT 3
#define VCT_MLP_PARAMS (VCT_MLP_IN * VCT_MLP_HIDDEN + VCT_MLP_HIDDEN + \
                        VCT_MLP_HIDDEN * VCT_MLP_HIDDEN + VCT_MLP_HIDDEN + \
                        VCT_MLP_HIDDEN * VCT_MLP_OUT + VCT_MLP_OUT)
/* Flat FP32 parameter vector: W1[64][34], b1[64], W2[64][64], b2[64],
 * W3[3][64], b3[3].  He-uniform initialisation from the seed. */
int vct_mlp_init(vct_ctx ctx, uint64_t seed);
int vct_mlp_set_params(vct_ctx ctx, const float* params, uint32_t n);
int vct_mlp_get_params(vct_ctx ctx, float* params, uint32_t n);
/* Network output for n feature rows (n x 34 floats, rounded to BF16 on the
 * device) -> n x 3 floats. */
int vct_mlp_forward(vct_ctx ctx, const float* features, uint32_t n, float* out);
/* vct_mlp_forward on device buffers, allocating nothing.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  features (n x 34 floats) and out (n x 3 floats) are
 * device pointers; the work is queued on `stream` (a hipStream_t; NULL: the
 * context's stream), after vct_mlp_set_params / vct_mlp_init has completed on
 * it (vct_sync).  The BF16 rows are staged in a fixed per-GPU array of
 * VCT_MLP_NOALLOC_MAX_ROWS rows (32 MiB), held until the process ends: larger
 * n is rejected with VCT_ERR_LIMIT.  Every call on a GPU shares that array, so
 * calls that overlap on one GPU race.  Outputs equal vct_mlp_forward's. */
#define VCT_MLP_NOALLOC_MAX_ROWS 262144
int vct_mlp_forward_noalloc(vct_ctx ctx, const float* features, uint32_t n, float* out, void* stream);
/* Loss mean((albedo * (y - target))^2) over n samples and 3 channels and its
 * gradient w.r.t. every parameter (the training kernels, no update). */
int vct_mlp_loss_grad(vct_ctx ctx, const float* features, const float* albedo, const float* target, uint32_t n,
                      float* grad, float* loss);
/* Training data from the current frame (after vct_voxelize, vct_mip_aniso
 * and vct_gbuffer): traces the full-resolution albedo-demodulated target and
 * the quarter-resolution input, then appends nsamples random pixels
 * (features, albedo, target) to the device-resident training set. */
int vct_upsample_collect(vct_ctx ctx, uint32_t nsamples, uint64_t seed);
int vct_upsample_clear_samples(vct_ctx ctx);
/* Replaces the training set with n host samples (features n x 34, albedo and
 * target n x 3). */
int vct_upsample_set_samples(vct_ctx ctx, const float* features, const float* albedo, const float* target,
                             uint32_t n);
int vct_upsample_sample_count(vct_ctx ctx, uint64_t* n);
/* Adam steps (beta 0.9 / 0.999) on minibatches of `batch` samples: step s
 * takes samples s*batch .. s*batch + batch - 1 (mod n) of a random order of
 * the n samples, drawn from `seed`. */
int vct_upsample_train(vct_ctx ctx, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* last_loss);
/* The learned form of Renderer::render: voxelize, mip, full-resolution
 * G-buffer, quarter-resolution cone trace, MLP upsampling (surface-guided:
 * a pixel with no quarter-resolution sample of its own surface nearby is
 * traced at full resolution instead), then the per-tile fallback: every
 * 32x32 tile is compared with a full-resolution trace at 16 probe pixels,
 * which estimates its error, and the tiles with the largest estimated errors
 * are traced again in full until the error left in the others is within a
 * share of the frame's budget for the bound (0.55 - 1.7/sqrt(covered tiles),
 * at least 0.25: margin for the estimate's noise; with the half-resolution
 * refinement on, every tile whose own estimate is below the bound is refined
 * instead).  Pass times [voxelize, mip, gbuffer, low-res trace, upsample,
 * fallback]. */
int vct_render_learned(vct_ctx ctx);
/* The PSNR bound in dB (peak 1, RGB over covered pixels, against the plain
 * full-resolution frame); default 40; <= 0 turns the fallback off. */
int vct_upsample_set_fallback(vct_ctx ctx, float min_tile_psnr_db);
/* Tiles of the viewport and tiles traced in full by the last learned frame. */
int vct_upsample_fallback_stats(vct_ctx ctx, uint32_t* tiles, uint32_t* fallback_tiles);
/* 0 (default): the frame budget above, picked tiles traced in full.  1:
 * every tile whose estimated PSNR is below the bound first has its odd-odd
 * pixels traced (half resolution) and interpolated from with the same
 * surface guidance, is checked again at 16 even-even probes, and is traced in
 * full only if it fails again. */
int vct_upsample_set_refine(vct_ctx ctx, int half_resolution);
/* Counters of the last learned frame: [tiles, tiles refined at half
 * resolution (or traced in full when that stage is off), tiles traced in
 * full, pixels traced after the quarter-resolution stage, pixels traced
 * after the half-resolution stage, probe pixels traced]. */
int vct_upsample_stats(vct_ctx ctx, uint32_t out[6]);
/* Pixels of the last learned frame traced at full resolution because no
 * quarter-resolution sample of their surface was near. */
int vct_upsample_traced_pixels(vct_ctx ctx, uint32_t* pixels);
int vct_upsample_lowres_download(vct_ctx ctx, float* rgba, size_t pixels);

#ifdef __cplusplus
}
#endif
#endif /* VCT_H */
