// vct_internal.h - shared definitions of the gfx950 voxel cone tracing core.
//
// Volume layout in HBM (all levels, all six direction maps, one allocation):
//   level 0 (res^3) is shared by the six direction maps: the reference writes
//   the same value into all six level-0 images (voxelize.frag:159-160), so one
//   copy is kept.  Levels 1..L-1 exist once per direction (mipmap.comp:4-5).
//   Each level is stored along a 3D Morton (Z-order) curve: every aligned
//   2^k cube of texels is one contiguous run (4x4x4 bricks are 64 elements,
//   256 B to 1 KiB), so a 2x2x2 trilinear footprint usually touches one or two
//   128-byte lines, and the eight texels under a mip texel m are elements
//   [8m, 8m + 8): the mip builder streams its input and output contiguously.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/vct.h"

#define VCT_HD __host__ __device__ __forceinline__

namespace vct {

// ---- wave-aggregated counters ----------------------------------------------
#ifdef __HIPCC__
// Adds 1 to counter[key] for every active lane, one atomic per distinct key
// in the wave; returns the lane's slot (the counter's old value + its rank
// among the lanes with the same key).
__device__ __forceinline__ uint32_t wave_add_one(uint32_t* counter, uint32_t key, bool active) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(active);
  // group the active lanes by key (ALU only), then one atomic per group, all
  // groups' atomics in one instruction: a single memory round trip
  int my_leader = lane;
  uint32_t my_rank = 0, my_count = 0;
  while (pending) {
    const int leader = __ffsll(pending) - 1;
    const uint32_t k = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(active && key == k);
    if (active && key == k) {
      my_leader = leader;
      my_rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    }
    if (lane == leader) my_count = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  uint32_t old = 0;
  if (active && lane == my_leader) old = atomicAdd(counter + key, my_count);
  old = __shfl(old, my_leader, 64);
  return old + my_rank;
}
#endif

// ---- Morton / brick addressing -------------------------------------------
VCT_HD uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
VCT_HD uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x3ffu;
  return v;
}
VCT_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
// Element index of texel (x,y,z) in a level: the 3D Morton code.  Every
// aligned 2^k cube is a contiguous run; in particular the 2x2x2 block under a
// mip texel m is elements [8m, 8m + 8) and any 4x4x4 brick is 64 elements.
VCT_HD uint32_t brick_index(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
// Inverse of brick_index.
VCT_HD void brick_coords(uint32_t idx, uint32_t& x, uint32_t& y, uint32_t& z) {
  x = compact3(idx);
  y = compact3(idx >> 1);
  z = compact3(idx >> 2);
}
VCT_HD uint64_t level_elems(uint32_t s) { return (uint64_t)s * s * s; }

// ---- storage codecs -------------------------------------------------------
VCT_HD uint32_t dtype_bytes(int dt) {
  return dt == VCT_RGBA8 ? 4u : (dt == VCT_FP32 ? 16u : 8u);
}

// One entry of the marcher's x-paired copy: texel a and its +x neighbour b.
// 16-bit storage pairs channel by channel (four words of (c(a), c(b))), so a
// two-term dot product weights both x-corners of a channel at once; RGBA8
// texels pair as two words.
VCT_HD uint4 pair_pack(uint2 a, uint2 b) {
  return make_uint4((a.x & 0xffffu) | (b.x << 16), (a.x >> 16) | (b.x & 0xffff0000u), (a.y & 0xffffu) | (b.y << 16),
                    (a.y >> 16) | (b.y & 0xffff0000u));
}
VCT_HD uint2 pair_pack(uint32_t a, uint32_t b) { return make_uint2(a, b); }

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).
VCT_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
VCT_HD float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

VCT_HD uint16_t f32_to_f16(float f) {
  _Float16 h = (_Float16)f;  // round to nearest even (v_cvt_f16_f32)
  return __builtin_bit_cast(uint16_t, h);
}
VCT_HD float f16_to_f32(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// GL unorm8 -> float: c / 255 (the GL definition; correctly rounded divide).
VCT_HD float unorm8_exact(uint32_t c) { return (float)c / 255.0f; }
// float -> GL unorm8: clamp to [0,1], scale by 255, round to nearest even.
VCT_HD uint32_t to_unorm8(float f) {
  f = f > 0.0f ? f : 0.0f;  // also maps NaN to 0
  f = f < 1.0f ? f : 1.0f;
  return (uint32_t)__builtin_rintf(f * 255.0f);
}

// ---- device material ---------------------------------------------------
// 64-byte record derived from vct_material (the fields the three shaders
// read: voxelize.frag:10-31, voxel_cone_tracing.frag:3-24).
struct DevMaterial {
  float diffuse[3];
  float shininess;
  float specular[3];
  float ior;
  float emission[3];
  float dissolve;
  float transmittance[3];
  int32_t illum;
};
static_assert(sizeof(DevMaterial) == 64, "DevMaterial must be 64 bytes");

VCT_HD bool illum_transmissive(int illum) {
  return illum == 4 || illum == 6 || illum == 7 || illum == 9;
}

struct DevLight {
  float pos[3];
  float color[3];
  float intensity;
  int32_t kind;
};

struct LightSet {
  DevLight l[VCT_MAX_LIGHTS];
  int32_t count;
};

}  // namespace vct

// ---- host side context --------------------------------------------------
// A model's triangles are stored on the device in Morton order of their
// centroids (spatially coherent rasterisation and voxel scatter); d_perm maps
// a stored position to the triangle's original index and d_inv back.  Every
// order-dependent rule (GL draw order for depth ties, the canonical order of
// the running average) uses the original index.
struct VctModel {
  uint32_t ntri = 0;
  float* d_pos = nullptr;     // 9 floats per triangle (stored order)
  float* d_nrm = nullptr;     // 9 floats per triangle (stored order)
  uint32_t* d_mat = nullptr;  // material id per triangle (stored order)
  uint32_t* d_perm = nullptr;  // stored -> original index
  uint32_t* d_inv = nullptr;   // original -> stored index
  std::vector<uint32_t> perm;
  std::vector<uint32_t> h_mat;
  float xform[16];
  float dims[3];
};

struct VctTex3D {
  bool alive = false;
  int w = 0, h = 0, d = 0, levels = 0, dtype = 0;
  std::vector<size_t> off;  // byte offset of each level (linear layout)
  std::vector<int> lw, lh, ld;
  void* dptr = nullptr;
  size_t bytes = 0;
};

struct VctBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// A per-frame table (the draws of a pass) copied to the device on a stream:
// pinned host staging and the device copy.  The staging area is rewritten
// only once the copy that last read it has completed (its event).
struct VctUpload {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  VctBuf dev;
};

struct vct_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;          // second stream (default priority): the G-buffer beside voxelize + mip
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // pinned host words the list sizes are copied into: [0] G-buffer
  // (triangle, tile) pairs, [1] G-buffer fan triangles of clipped triangles,
  // [2] voxelize (triangle, brick) pairs
  uint32_t* h_readback = nullptr;
  int width = 0, height = 0;
  unsigned res = 128;
  int levels = 7;
  int dtype = VCT_RGBA8;
  int average = VCT_AVG_REFERENCE;
  float cube_size = 1.0f;
  float view[16], proj[16];
  bool en_direct = true, en_diffuse = true, en_specular = true, en_shadow = true;
  int view_voxel_dir = 0;
  float view_voxel_lod = 0.0f;
  int cone_set = VCT_CONES_REF9;

  std::vector<vct_material> materials;
  VctBuf d_materials;
  bool materials_dirty = true;

  std::vector<VctModel> models;
  std::vector<uint32_t> queue;
  std::vector<vct_light> lights;

  // radiance volume
  VctBuf volume;
  bool volume_valid = false;
  uint64_t lvl_off[16][6];  // byte offsets
  uint32_t lvl_size[16];
  uint64_t volume_version = 1;  // advanced by every write of the volume
  // x-paired copy of the volume for the cone marcher (16-bit and RGBA8 storage)
  VctBuf pairs;
  uint64_t pair_off[16][6];
  uint64_t pairs_version = 0;  // volume_version the copy was built from

  // voxelize scratch
  VctBuf tri_count, tri_off, tri_base;  // per-brick triangle counts + cursors, triangle lists, per-brick offsets
  VctBuf vox_box;                      // voxelize: packed brick box per triangle
  VctBuf frags_sorted, vox_count, vox_off, scan_tmp, g_scan_tmp, counters;
  VctBuf setup;  // per-frame triangle list (model index, transform)
  // voxelize: the material terms of the queued models' triangles, in frame
  // order; rebuilt only when shade_epoch (bumped by every material change and
  // model material assignment) or the queued models differ from those it
  // was built for
  // vct_frame_front also launches the quarter-resolution G-buffer gather on
  // the G-buffer's stream, right after the G-buffer (the learned frame)
  bool front_lowres = false;
  VctBuf vox_shade;
  uint64_t shade_epoch = 0, vox_shade_epoch = ~0ull;
  std::vector<uint32_t> vox_shade_models;
  void* vox_shade_ptr = nullptr;
  uint64_t last_fragments = 0, last_candidates = 0;
  // Exact-mean voxelize lists sized without a host wait: once a total is
  // known for the queued triangle count, the list holds vox_cap entries and a
  // brick whose range runs past them walks every triangle's brick box
  // instead.  Each frame's (triangle, brick) pair count comes back into
  // h_readback[3] asynchronously, complete once vox_total_ev is.
  uint32_t vox_known_total = 0, vox_known_ntri = 0, vox_cap = 0;
  hipEvent_t vox_total_ev = nullptr;
  bool vox_total_pending = false;
  bool vox_sized = false;            // this frame's list was launched against vox_cap
  bool vox_pairs_on_device = false;  // last_candidates still to be read from counters[2]

  // G-buffer + image
  VctBuf g_list, g_pos, g_nrm, image;
  VctBuf g_count, g_off, g_recs;  // per-tile counts + cursors, per-tile offsets, per-triangle records
  VctUpload g_draws;               // the G-buffer's draw table
  VctUpload v_draws;               // voxelize's draw table
  VctBuf g_fans;                   // records and resolve data of clipped triangles' fan triangles
  VctBuf g_clip;                   // the frame's triangles that need clipping
  uint64_t last_gbuffer_refs = 0;  // (triangle, tile) pairs of the last G-buffer
  // what the last rasterised G-buffer left for the voxel view's in-order
  // composite (its tile lists, records and draw table); invalid after an
  // uploaded G-buffer
  struct {
    bool valid = false;
    int W = 0, H = 0, ndraw = 0;
    uint32_t ntri = 0, nfans = 0;
    uint64_t total = 0;
    float PV[16];
  } g_frame;
  VctBuf vv_keys, vv_vals, vv_ends, vv_temp;  // the voxel view's draw-order sort of the tile lists

  std::vector<VctTex3D> textures;

  // learned upsampler
  VctBuf mlp_params, mlp_m1, mlp_m2, mlp_packed, mlp_bias, mlp_grad, mlp_partial;
  bool mlp_ready = false;
  int mlp_step = 0;
  VctBuf low_gpos, low_gnrm, low_img, full_target, sample_pix;
  VctBuf train_X, train_A, train_Y, train_valid;
  VctBuf train_perm, batch_X, batch_A, batch_Y;  // minibatch order and the gathered minibatch
  size_t train_n = 0, train_cap = 0;
  // per-tile fallback of the learned frame
  VctBuf probe_pix, probe_val, fb_list;  // fb_list: [count][tiles...]
  VctBuf fb_pix;                          // [count][pixels...] traced at full resolution
  VctBuf fb_list2, fb_pix2, fb_flag;      // tiles refined again, pixels listed by the half-res pass, tile states
  VctBuf half_img, probe_pix2, probe_val2;
  VctBuf fb_est;                          // per-tile error estimate + covered probes
  bool fb_half = false;                   // per-tile bound + half-resolution stage instead of the frame budget
  bool fb_half_ran = false;
  int probe_w = 0, probe_h = 0;
  float fb_psnr_db = 40.0f;  // minimum estimated tile PSNR; <= 0 disables the fallback
  bool fb_ran = false;

  hipEvent_t ev[12];  // start / end of 6 passes
  float pass_ms[6] = {0, 0, 0, 0, 0, 0};
  bool ev_recorded[6] = {false, false, false, false, false, false};
  uint32_t time_mask = 0x3fu;  // passes timed with events (bit k: pass k; vct_set_pass_timing)
  hipEvent_t g_count_ev = nullptr;    // the G-buffer's list sizes copied to h_readback[0..1]
};

// ---- internal helpers (vct_runtime.hip) ---------------------------------
int vct_set_error(int code, const std::string& msg);
int vct_hip_check(hipError_t e, const char* what);
int vct_ensure(VctBuf& b, size_t bytes, vct_context* ctx);
// Copies bytes of a host table to u.dev on ctx->stream (asynchronously).
int vct_upload(vct_context* ctx, VctUpload& u, const void* data, size_t bytes);
void vct_upload_free(VctUpload& u);
#define VCT_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return vct_hip_check(_e, #call); \
  } while (0)

// Records the start (end = false) or the end event of pass `slot` on stream
// s when pass timing covers that pass; an untimed pass reads 0 ms.
inline int vct_pass_event(vct_context* ctx, int slot, bool end, hipStream_t s) {
  if (!((ctx->time_mask >> slot) & 1u)) {
    ctx->ev_recorded[slot] = false;
    ctx->pass_ms[slot] = 0.0f;
    return VCT_OK;
  }
  VCT_HIP(hipEventRecord(ctx->ev[2 * slot + (end ? 1 : 0)], s));
  if (end) ctx->ev_recorded[slot] = true;
  return VCT_OK;
}

// kernels launchers
int vct_launch_mip(vct_context* ctx);
int vct_ensure_pairs(vct_context* ctx);  // x-paired copy of the volume for the marcher
// Offsets of the paired copy's maps (ctx->pair_off) and its buffer.
int vct_pair_layout(vct_context* ctx);
// The paired copy of levels first.. (the mip kernel writes levels 0 and 1).
int vct_pairs_from(vct_context* ctx, int first);
int vct_launch_voxel_view(vct_context* ctx);  // the voxel view blended in draw order (last G-buffer's lists)
int vct_frame_front(vct_context* ctx);   // voxelize + mip beside the G-buffer, joined
int vct_launch_voxelize(vct_context* ctx);
int vct_launch_gbuffer(vct_context* ctx);
// The two passes that size a triangle list on the device, in two phases:
// phase 1 launches the work up to the list size and its copy to pinned host
// memory, phase 2 waits for that copy and launches the rest (3 = both).  The
// frame front runs phase 1 of both passes before either waits.
int vct_launch_voxelize_phase(vct_context* ctx, int phase);
int vct_launch_gbuffer_phase(vct_context* ctx, int phase);
// The quarter-resolution G-buffer of the learned frame (its buffers already
// sized), on the context stream.
int vct_launch_lowres_gather(vct_context* ctx);
int vct_launch_trace(vct_context* ctx);
// Trace of a W x H G-buffer into out; demodulate: final / max(Kd, eps) with
// the pixel's linear depth in w (the learned path's buffers).
// Pixel selection of a trace launch (zero-initialised: every pixel).
struct TraceSelect {
  const uint32_t* pix_list;    // device list of pixels (y * W + x); output k = pixel k
  uint32_t npix_list;          // count, or the bound of *pix_count (grid size)
  const uint32_t* pix_count;   // optional device-side count
  int pix_scatter;             // 1: output indexed like the G-buffer
  const uint32_t* tile_list;   // device list of 32x32 tiles; output indexed like the G-buffer
  const uint32_t* tile_count;  // device-side count of listed tiles
  uint32_t max_tiles;          // grid bound (host side)
  int tiles_x;
  const uint32_t* tile_flag;   // pixel lists: trace only where tile_flag[tile] == flag_want
  uint32_t flag_want;
  int grid_stride;             // tile lists: 2 = the odd-odd pixels, written to a half-resolution image
  int half_w;
};
int vct_trace_select(vct_context* ctx, const float4* gpos, const float4* gnrm, float4* out, int W, int H,
                     int demodulate, unsigned long long* step_counter, const TraceSelect& sel);
int vct_trace_into(vct_context* ctx, const float4* gpos, const float4* gnrm, float4* out, int W, int H,
                   int demodulate, unsigned long long* step_counter);
int vct_launch_scan_u32(vct_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                        uint32_t* total_dev, VctBuf* tmp = nullptr);
int vct_tex3d_mip_launch(vct_context* ctx, VctTex3D& t);
// Device address of the no-allocation entry point's staging array, taken once
// per device (the current one) at context creation.
int vct_noalloc_setup(int device);
int vct_tex3d_clear_launch(vct_context* ctx, VctTex3D& t, const float rgba[4]);
