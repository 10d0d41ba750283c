// vct_runtime.hip - context, resources and the C ABI of include/vct.h.
//
// Stands in for the resource side of Renderer (renderer.cpp:100-281,
// 392-405, 643-677) and for Texture3D (texture_3d.cpp:3-51).  All device
// work of a context is queued in order on one HIP stream created with the
// default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "vct_internal.h"

static thread_local std::string g_last_error;

int vct_set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int vct_hip_check(hipError_t e, const char* what) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return vct_set_error(e == hipErrorOutOfMemory ? VCT_ERR_OUT_OF_MEMORY : VCT_ERR_HIP, buf);
}

int vct_ensure(VctBuf& b, size_t bytes, vct_context* ctx) {
  if (bytes <= b.cap && b.p) return VCT_OK;
  if (b.p) {
    VCT_HIP(hipStreamSynchronize(ctx->stream));
    VCT_HIP(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    char msg[160];
    snprintf(msg, sizeof(msg), "device allocation of %zu bytes failed", want);
    return vct_set_error(VCT_ERR_OUT_OF_MEMORY, msg);
  }
  b.cap = want;
  return VCT_OK;
}

static void free_buf(VctBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

int vct_upload(vct_context* ctx, VctUpload& u, const void* data, size_t bytes) {
  if (!u.ev) VCT_HIP(hipEventCreateWithFlags(&u.ev, hipEventDisableTiming));
  else VCT_HIP(hipEventSynchronize(u.ev));  // the last copy out of the staging area has completed
  if (bytes > u.cap) {
    if (u.host) VCT_HIP(hipHostFree(u.host));
    u.host = nullptr;
    u.cap = 0;
    const size_t want = bytes < 4096 ? 4096 : bytes;
    VCT_HIP(hipHostMalloc(&u.host, want, hipHostMallocDefault));
    u.cap = want;
  }
  int rc = vct_ensure(u.dev, bytes, ctx);
  if (rc) return rc;
  memcpy(u.host, data, bytes);
  VCT_HIP(hipMemcpyAsync(u.dev.p, u.host, bytes, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipEventRecord(u.ev, ctx->stream));
  return VCT_OK;
}

void vct_upload_free(VctUpload& u) {
  if (u.ev) {
    (void)hipEventSynchronize(u.ev);
    (void)hipEventDestroy(u.ev);
  }
  if (u.host) (void)hipHostFree(u.host);
  free_buf(u.dev);
  u.ev = nullptr;
  u.host = nullptr;
  u.cap = 0;
}

// ---- scan (exclusive, u32) --------------------------------------------------
// Three-phase scan: per-block reduce, scan of block sums (recursive), per-block
// scan plus block offset.  1024 elements per 256-thread block.
namespace {
constexpr int kScanBlock = 1024;

__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* sh, uint32_t& total) {
  // wave-level inclusive scan then combine the four waves
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += sh[w];
  total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return off + x - v;
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(const uint32_t* in, uint32_t n, uint32_t* sums) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * kScanBlock;
  uint32_t s = 0;
  for (int k = 0; k < 4; ++k) {
    uint32_t i = base + threadIdx.x * 4 + k;
    if (i < n) s += in[i];
  }
  uint32_t total;
  block_exclusive_scan_256(s, sh, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const uint32_t* in, uint32_t n, const uint32_t* block_off,
                                                        uint32_t* out, uint32_t* total_out) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * kScanBlock;
  uint32_t v[4];
  uint32_t s = 0;
  for (int k = 0; k < 4; ++k) {
    uint32_t i = base + threadIdx.x * 4 + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t ex = block_exclusive_scan_256(s, sh, total);
  uint32_t run = ex + (block_off ? block_off[blockIdx.x] : 0u);
  for (int k = 0; k < 4; ++k) {
    uint32_t i = base + threadIdx.x * 4 + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) *total_out = run;
}
}  // namespace

// Exclusive scan of n u32 values; *total_dev (device) receives the sum.
// Scratch comes from *tmp, by default ctx->scan_tmp (scans that may run at
// the same time on two streams need scratch of their own).  Sums must fit in
// 32 bits.
int vct_launch_scan_u32(vct_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev,
                        VctBuf* tmp_buf) {
  VctBuf& scratch = tmp_buf ? *tmp_buf : ctx->scan_tmp;
  if (n == 0) {
    if (total_dev) VCT_HIP(hipMemsetAsync(total_dev, 0, 4, ctx->stream));
    return VCT_OK;
  }
  // level sizes
  std::vector<uint32_t> sizes;
  uint32_t m = n;
  while (true) {
    sizes.push_back(m);
    if (m <= (uint32_t)kScanBlock) break;
    m = (m + kScanBlock - 1) / kScanBlock;
  }
  // scratch: for each level >0 a sums array and an offsets array
  size_t need = 0;
  for (size_t i = 1; i < sizes.size(); ++i) need += 2 * ((size_t)sizes[i] + 64);
  int rc = vct_ensure(scratch, need * 4 + 256, ctx);
  if (rc) return rc;
  uint32_t* tmp = reinterpret_cast<uint32_t*>(scratch.p);
  std::vector<uint32_t*> sums(sizes.size(), nullptr), offs(sizes.size(), nullptr);
  size_t pos = 0;
  for (size_t i = 1; i < sizes.size(); ++i) {
    sums[i] = tmp + pos;
    pos += sizes[i] + 64;
    offs[i] = tmp + pos;
    pos += sizes[i] + 64;
  }
  // reduce down
  const uint32_t* cur = in;
  for (size_t i = 0; i + 1 < sizes.size(); ++i) {
    uint32_t nb = (sizes[i] + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(256), 0, ctx->stream, cur, sizes[i], sums[i + 1]);
    cur = sums[i + 1];
  }
  // scan the top level in one block, then apply downwards
  for (size_t i = sizes.size(); i-- > 0;) {
    const uint32_t* src = (i == 0) ? in : sums[i];
    uint32_t* dst = (i == 0) ? out : offs[i];
    const uint32_t* boff = (i + 1 < sizes.size()) ? offs[i + 1] : nullptr;
    uint32_t nb = (sizes[i] + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(256), 0, ctx->stream, src, sizes[i], boff, dst,
                       i == 0 ? total_dev : nullptr);
  }
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// ---- volume ------------------------------------------------------------------
// Byte layout of the volume: level 0 (one map shared by all directions),
// then 6 directional maps per level, each 256-byte aligned.  Returns the
// total size; fills off[l][d] when it is given.
static uint64_t volume_layout(unsigned res, int levels, int dtype, uint64_t (*off)[6]) {
  const uint32_t bpe = vct::dtype_bytes(dtype);
  uint64_t total = 0;
  for (int l = 0; l < levels; ++l) {
    const uint64_t bytes = (vct::level_elems(res >> l) * bpe + 255) & ~255ull;
    for (int d = 0; d < 6; ++d) {
      if (off) off[l][d] = total;
      if (l > 0) total += bytes;
    }
    if (l == 0) total += bytes;
  }
  return total;
}

static int alloc_volume(vct_context* ctx) {
  for (int l = 0; l < ctx->levels; ++l) ctx->lvl_size[l] = ctx->res >> l;
  const uint64_t off = volume_layout(ctx->res, ctx->levels, ctx->dtype, ctx->lvl_off);
  ++ctx->volume_version;
  int rc = vct_ensure(ctx->volume, off, ctx);
  if (rc) return rc;
  VCT_HIP(hipMemsetAsync(ctx->volume.p, 0, off, ctx->stream));
  ctx->volume_valid = true;
  return VCT_OK;
}

static int validate_grid(unsigned res, int levels, int dtype, int average) {
  if (res < 4 || res > 1024 || (res & (res - 1)))
    return vct_set_error(VCT_ERR_INVALID_ARG, "grid resolution must be a power of two in [4, 1024]");
  int maxl = 1;
  for (unsigned r = res; r > 1; r >>= 1) ++maxl;
  if (levels < 1 || levels > maxl || levels > 16)
    return vct_set_error(VCT_ERR_INVALID_ARG, "levels must be in [1, log2(res)+1]");
  if (dtype < VCT_RGBA8 || dtype > VCT_FP32) return vct_set_error(VCT_ERR_INVALID_ARG, "bad storage dtype");
  if (average != VCT_AVG_REFERENCE && average != VCT_AVG_EXACT)
    return vct_set_error(VCT_ERR_INVALID_ARG, "bad averaging mode");
  if (average == VCT_AVG_REFERENCE && dtype != VCT_RGBA8)
    return vct_set_error(VCT_ERR_INVALID_ARG, "the reference running average needs RGBA8 storage");
  // the cone marcher addresses texels by 32-bit byte offsets (512^3 with FP32
  // or with 16-bit storage, whose x-paired copy is twice the volume, is the
  // largest grid that fits)
  const uint64_t paired = dtype == VCT_FP32 ? 1 : 2;  // the cone marcher's x-paired copy
  if (paired * volume_layout(res, levels, dtype, nullptr) > 0xffffffffull)
    return vct_set_error(VCT_ERR_INVALID_ARG, "radiance volume above 4 GiB (use a smaller grid or 16-bit storage)");
  return VCT_OK;
}

static vct::DevMaterial to_dev(const vct_material& m) {
  vct::DevMaterial d;
  for (int i = 0; i < 3; ++i) {
    d.diffuse[i] = m.diffuse[i];
    d.specular[i] = m.specular[i];
    d.emission[i] = m.emission[i];
    d.transmittance[i] = m.transmittance[i];
  }
  d.shininess = m.shininess;
  d.ior = m.ior;
  d.dissolve = m.dissolve;
  d.illum = m.illum;
  return d;
}

static int sync_materials(vct_context* ctx) {
  if (!ctx->materials_dirty) return VCT_OK;
  size_t n = ctx->materials.size();
  std::vector<vct::DevMaterial> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = to_dev(ctx->materials[i]);
  int rc = vct_ensure(ctx->d_materials, n * sizeof(vct::DevMaterial), ctx);
  if (rc) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->d_materials.p, h.data(), n * sizeof(vct::DevMaterial), hipMemcpyHostToDevice,
                         ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->materials_dirty = false;
  return VCT_OK;
}

static void identity(float* m) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

#define CTX_CHECK(ctx)                                                          \
  do {                                                                          \
    if (!(ctx)) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");       \
    hipError_t _se = hipSetDevice((ctx)->device);                               \
    if (_se != hipSuccess) return vct_hip_check(_se, "hipSetDevice");           \
  } while (0)

// The passes before the trace of a frame: the G-buffer on the second stream
// beside voxelize and mip on the context stream (they read the same geometry
// and write disjoint buffers; each pass is timed on its own stream, so the
// pass times overlap); the context stream then waits for the G-buffer.
// Voxelize sizes its triangle list from earlier frames' counts and is
// launched whole without a host wait (after the first frame, which reads its
// count back); the G-buffer reads its list size and fan count back, the host
// waiting on the copy's event.  Starting the G-buffer's raster only after
// voxelize's brick kernel, beside the HBM-bound mip, measured slower: its
// long workgroups take the slots the mip's 512-lane ones need.
int vct_frame_front(vct_context* ctx) {
  int rc;
  if ((rc = sync_materials(ctx))) return rc;
  // runs fn with the side stream as the context stream
  auto on_side = [ctx](auto&& fn) {
    std::swap(ctx->stream, ctx->side);
    const int r = fn();
    std::swap(ctx->stream, ctx->side);
    return r;
  };
  VCT_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
  VCT_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  if ((rc = vct_pass_event(ctx, 2, false, ctx->side))) return rc;
  if ((rc = on_side([ctx] { return vct_launch_gbuffer_phase(ctx, 1); }))) return rc;
  ++ctx->volume_version;
  if ((rc = vct_pass_event(ctx, 0, false, ctx->stream))) return rc;
  if ((rc = vct_launch_voxelize_phase(ctx, 1))) return rc;
  if ((rc = on_side([ctx] { return vct_launch_gbuffer_phase(ctx, 2); }))) return rc;
  if (ctx->front_lowres && (rc = on_side([ctx] { return vct_launch_lowres_gather(ctx); }))) return rc;
  if ((rc = vct_pass_event(ctx, 2, true, ctx->side))) return rc;
  if ((rc = vct_launch_voxelize_phase(ctx, 2))) return rc;
  if ((rc = vct_pass_event(ctx, 0, true, ctx->stream))) return rc;
  VCT_HIP(hipEventRecord(ctx->ev_join, ctx->side));
  if ((rc = vct_mip_aniso(ctx))) return rc;
  VCT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  return VCT_OK;
}

extern "C" {

const char* vct_last_error(void) { return g_last_error.c_str(); }

int vct_device_count(int* out) {
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *out = n;
  return VCT_OK;
}

int vct_create(int width, int height, vct_ctx* out) {
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  *out = nullptr;
  if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
    return vct_set_error(VCT_ERR_INVALID_ARG, "viewport must be in [1, 16384]^2");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return vct_set_error(VCT_ERR_NO_DEVICE, "no HIP device is visible");
  }
  int dev = 0;
  VCT_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  VCT_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    return vct_set_error(VCT_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                                ", this library is built for gfx950 only");
  }
  vct_context* ctx = new (std::nothrow) vct_context();
  if (!ctx) return vct_set_error(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
  ctx->device = dev;
  ctx->width = width;
  ctx->height = height;
  identity(ctx->view);
  identity(ctx->proj);
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
  if (e != hipSuccess) {
    vct_destroy(ctx);
    return vct_hip_check(e, "hipStreamCreate");
  }
  for (int i = 0; i < 12 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->ev[i]);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->vox_total_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->g_count_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    vct_destroy(ctx);
    return vct_hip_check(e, "hipEventCreate");
  }
  e = hipHostMalloc((void**)&ctx->h_readback, 64, hipHostMallocDefault);
  if (e != hipSuccess) {
    ctx->h_readback = nullptr;
    vct_destroy(ctx);
    return vct_hip_check(e, "hipHostMalloc");
  }
  // default material (renderer.cpp:85-98,115-128)
  vct_material dm;
  memset(&dm, 0, sizeof(dm));
  dm.ambient[0] = 1;
  dm.ambient[2] = 1;
  dm.diffuse[0] = 1;
  dm.diffuse[2] = 1;
  dm.shininess = 1;
  dm.ior = 1;
  dm.dissolve = 1;
  dm.illum = 0;
  ctx->materials.push_back(dm);
  int rc = alloc_volume(ctx);
  if (!rc) rc = vct_noalloc_setup(dev);  // the staging array of vct_mlp_forward_noalloc
  if (rc) {
    vct_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return VCT_OK;
}

int vct_destroy(vct_ctx ctx) {
  if (!ctx) return VCT_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& m : ctx->models) {
    if (m.d_pos) (void)hipFree(m.d_pos);
    if (m.d_nrm) (void)hipFree(m.d_nrm);
    if (m.d_mat) (void)hipFree(m.d_mat);
    if (m.d_perm) (void)hipFree(m.d_perm);
    if (m.d_inv) (void)hipFree(m.d_inv);
  }
  for (auto& t : ctx->textures)
    if (t.dptr) (void)hipFree(t.dptr);
  VctBuf* bufs[] = {&ctx->d_materials, &ctx->volume, &ctx->tri_count, &ctx->tri_off,
                    &ctx->tri_base, &ctx->frags_sorted, &ctx->vox_count, &ctx->vox_off,
                    &ctx->scan_tmp, &ctx->g_scan_tmp, &ctx->counters, &ctx->setup, &ctx->g_list, &ctx->g_pos,
                    &ctx->g_nrm, &ctx->image, &ctx->g_count, &ctx->g_off, &ctx->g_recs,
                    &ctx->mlp_params, &ctx->mlp_m1, &ctx->mlp_m2, &ctx->mlp_packed, &ctx->mlp_bias,
                    &ctx->mlp_grad, &ctx->mlp_partial, &ctx->low_gpos, &ctx->low_gnrm, &ctx->low_img,
                    &ctx->full_target, &ctx->sample_pix, &ctx->train_X, &ctx->train_A, &ctx->train_Y,
                    &ctx->train_valid, &ctx->probe_pix, &ctx->probe_val, &ctx->fb_list, &ctx->fb_pix,
                    &ctx->fb_list2, &ctx->fb_pix2, &ctx->fb_flag, &ctx->half_img, &ctx->probe_pix2,
                    &ctx->probe_val2, &ctx->pairs, &ctx->fb_est, &ctx->g_fans, &ctx->g_clip,
                    &ctx->train_perm, &ctx->batch_X, &ctx->batch_A, &ctx->batch_Y, &ctx->vox_box, &ctx->vox_shade,
                    &ctx->vv_keys, &ctx->vv_vals, &ctx->vv_ends, &ctx->vv_temp};
  for (VctBuf* b : bufs) free_buf(*b);
  vct_upload_free(ctx->g_draws);
  vct_upload_free(ctx->v_draws);
  for (int i = 0; i < 12; ++i)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->vox_total_ev) (void)hipEventDestroy(ctx->vox_total_ev);
  if (ctx->g_count_ev) (void)hipEventDestroy(ctx->g_count_ev);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->h_readback) (void)hipHostFree(ctx->h_readback);
  delete ctx;
  return VCT_OK;
}

int vct_sync(vct_ctx ctx) {
  CTX_CHECK(ctx);
  VCT_HIP(hipStreamSynchronize(ctx->side));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_set_viewport(vct_ctx ctx, int width, int height) {
  CTX_CHECK(ctx);
  if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
    return vct_set_error(VCT_ERR_INVALID_ARG, "viewport must be in [1, 16384]^2");
  ctx->width = width;
  ctx->height = height;
  return VCT_OK;
}

int vct_set_grid_format(vct_ctx ctx, unsigned res, int levels, int dtype, int average) {
  CTX_CHECK(ctx);
  int rc = validate_grid(res, levels, dtype, average);
  if (rc) return rc;
  if (res == ctx->res && levels == ctx->levels && dtype == ctx->dtype && average == ctx->average &&
      ctx->volume_valid) {
    return VCT_OK;
  }
  ctx->res = res;
  ctx->levels = levels;
  ctx->dtype = dtype;
  ctx->average = average;
  return alloc_volume(ctx);
}

int vct_set_grid_resolution(vct_ctx ctx, unsigned res) {
  CTX_CHECK(ctx);
  int levels = 7;  // create_tex_3d(res, res, res, 7), renderer.cpp:186
  int maxl = 1;
  for (unsigned r = res; r > 1; r >>= 1) ++maxl;
  if (levels > maxl) levels = maxl;
  return vct_set_grid_format(ctx, res, levels, ctx->dtype, ctx->average);
}

int vct_set_grid_size(vct_ctx ctx, float size) {
  CTX_CHECK(ctx);
  if (!(size > 0.0f) || !std::isfinite(size)) return vct_set_error(VCT_ERR_INVALID_ARG, "grid size must be > 0");
  ctx->cube_size = size;
  return VCT_OK;
}

int vct_get_grid(vct_ctx ctx, int* res, int* levels, int* dtype, int* average) {
  CTX_CHECK(ctx);
  if (res) *res = (int)ctx->res;
  if (levels) *levels = ctx->levels;
  if (dtype) *dtype = ctx->dtype;
  if (average) *average = ctx->average;
  return VCT_OK;
}

int vct_set_camera_transform(vct_ctx ctx, const float view[16], const float projection[16]) {
  CTX_CHECK(ctx);
  if (!view || !projection) return vct_set_error(VCT_ERR_INVALID_ARG, "null matrix");
  memcpy(ctx->view, view, 64);
  memcpy(ctx->proj, projection, 64);
  return VCT_OK;
}

int vct_add_material(vct_ctx ctx, const vct_material* m, uint32_t* out_id) {
  return vct_add_materials(ctx, m, 1, out_id);
}

int vct_add_materials(vct_ctx ctx, const vct_material* m, uint32_t n, uint32_t* first_id) {
  CTX_CHECK(ctx);
  if (!m || !first_id || n == 0) return vct_set_error(VCT_ERR_INVALID_ARG, "null material or n == 0");
  *first_id = (uint32_t)ctx->materials.size();
  ctx->materials.insert(ctx->materials.end(), m, m + n);
  ctx->materials_dirty = true;
  ++ctx->shade_epoch;
  return VCT_OK;
}

int vct_upload_material_data(vct_ctx ctx, const vct_material* m, uint32_t id) {
  CTX_CHECK(ctx);
  if (!m) return vct_set_error(VCT_ERR_INVALID_ARG, "null material");
  if (id >= ctx->materials.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "material id out of range");
  ctx->materials[id] = *m;
  ctx->materials_dirty = true;
  ++ctx->shade_epoch;
  return VCT_OK;
}

int vct_get_material(vct_ctx ctx, uint32_t id, vct_material* out) {
  CTX_CHECK(ctx);
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  if (id >= ctx->materials.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "material id out of range");
  *out = ctx->materials[id];
  return VCT_OK;
}

int vct_model_create(vct_ctx ctx, const float* positions, const float* normals, const uint32_t* tri_material,
                     uint32_t ntri, uint32_t* out_id) {
  CTX_CHECK(ctx);
  if (!out_id) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  *out_id = VCT_INVALID_ID;
  if (ntri > 0 && (!positions || !normals)) return vct_set_error(VCT_ERR_INVALID_ARG, "null geometry");
  if (ntri > (1u << 28)) return vct_set_error(VCT_ERR_LIMIT, "at most 2^28 triangles per model");
  VctModel m;
  m.ntri = ntri;
  identity(m.xform);
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < (uint64_t)ntri * 3; ++i)
    for (int k = 0; k < 3; ++k) {
      float v = positions[i * 3 + k];
      mn[k] = std::min(mn[k], v);
      mx[k] = std::max(mx[k], v);
    }
  for (int k = 0; k < 3; ++k) m.dims[k] = ntri ? mx[k] - mn[k] : 0.0f;
  m.h_mat.assign(ntri, 0u);
  if (tri_material) {
    for (uint32_t i = 0; i < ntri; ++i) {
      if (tri_material[i] >= ctx->materials.size())
        return vct_set_error(VCT_ERR_BAD_HANDLE, "triangle material id out of range");
      m.h_mat[i] = tri_material[i];
    }
  }
  size_t b9 = (size_t)ntri * 9 * sizeof(float);
  if (ntri) {
    // Morton order of the centroids (10 bits per axis over the centroid box)
    std::vector<std::pair<uint32_t, uint32_t>> key(ntri);
    float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> cen((size_t)ntri * 3);
    for (uint32_t t = 0; t < ntri; ++t)
      for (int k = 0; k < 3; ++k) {
        const float c = (positions[t * 9 + k] + positions[t * 9 + 3 + k] + positions[t * 9 + 6 + k]) / 3.0f;
        cen[(size_t)t * 3 + k] = c;
        cmn[k] = std::min(cmn[k], c);
        cmx[k] = std::max(cmx[k], c);
      }
    for (uint32_t t = 0; t < ntri; ++t) {
      uint32_t q[3];
      for (int k = 0; k < 3; ++k) {
        const float ext = cmx[k] - cmn[k];
        float f = ext > 0.0f ? (cen[(size_t)t * 3 + k] - cmn[k]) / ext : 0.0f;
        f = std::isfinite(f) ? std::min(std::max(f, 0.0f), 1.0f) : 0.0f;
        q[k] = (uint32_t)(f * 1023.0f);
      }
      key[t] = {vct::morton3(q[0], q[1], q[2]), t};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                       return a.first < b.first;
                     });
    m.perm.resize(ntri);
    std::vector<uint32_t> inv(ntri), mat_s(ntri);
    std::vector<float> pos_s(b9 / 4), nrm_s(b9 / 4);
    for (uint32_t s = 0; s < ntri; ++s) {
      const uint32_t o = key[s].second;
      m.perm[s] = o;
      inv[o] = s;
      mat_s[s] = m.h_mat[o];
      memcpy(&pos_s[(size_t)s * 9], positions + (size_t)o * 9, 36);
      memcpy(&nrm_s[(size_t)s * 9], normals + (size_t)o * 9, 36);
    }
    m.h_mat = mat_s;
    VCT_HIP(hipMalloc(&m.d_pos, b9));
    VCT_HIP(hipMalloc(&m.d_nrm, b9));
    VCT_HIP(hipMalloc(&m.d_mat, (size_t)ntri * 4));
    VCT_HIP(hipMalloc(&m.d_perm, (size_t)ntri * 4));
    VCT_HIP(hipMalloc(&m.d_inv, (size_t)ntri * 4));
    VCT_HIP(hipMemcpyAsync(m.d_pos, pos_s.data(), b9, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipMemcpyAsync(m.d_nrm, nrm_s.data(), b9, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipMemcpyAsync(m.d_mat, m.h_mat.data(), (size_t)ntri * 4, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipMemcpyAsync(m.d_perm, m.perm.data(), (size_t)ntri * 4, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipMemcpyAsync(m.d_inv, inv.data(), (size_t)ntri * 4, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
  }
  ctx->models.push_back(std::move(m));
  *out_id = (uint32_t)(ctx->models.size() - 1);
  return VCT_OK;
}

int vct_set_model_material(vct_ctx ctx, uint32_t material_id, uint32_t model_id) {
  CTX_CHECK(ctx);
  if (model_id >= ctx->models.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "model id out of range");
  if (material_id >= ctx->materials.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "material id out of range");
  VctModel& m = ctx->models[model_id];
  std::fill(m.h_mat.begin(), m.h_mat.end(), material_id);
  ++ctx->shade_epoch;
  if (m.ntri) {
    VCT_HIP(hipMemcpyAsync(m.d_mat, m.h_mat.data(), (size_t)m.ntri * 4, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return VCT_OK;
}

int vct_set_model_transform(vct_ctx ctx, uint32_t model_id, const float m[16]) {
  CTX_CHECK(ctx);
  if (!m) return vct_set_error(VCT_ERR_INVALID_ARG, "null matrix");
  if (model_id >= ctx->models.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "model id out of range");
  memcpy(ctx->models[model_id].xform, m, 64);
  return VCT_OK;
}

int vct_get_model_dimensions(vct_ctx ctx, uint32_t model_id, float out[3]) {
  CTX_CHECK(ctx);
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  if (model_id >= ctx->models.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "model id out of range");
  memcpy(out, ctx->models[model_id].dims, 12);
  return VCT_OK;
}

__attribute__((visibility("hidden"))) int vct_model_set_dims(vct_ctx ctx, uint32_t model_id, const float dims[3]) {
  CTX_CHECK(ctx);
  if (model_id >= ctx->models.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "model id out of range");
  memcpy(ctx->models[model_id].dims, dims, 12);
  return VCT_OK;
}

int vct_model_triangle_count(vct_ctx ctx, uint32_t model_id, uint32_t* out) {
  CTX_CHECK(ctx);
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  if (model_id >= ctx->models.size()) return vct_set_error(VCT_ERR_BAD_HANDLE, "model id out of range");
  *out = ctx->models[model_id].ntri;
  return VCT_OK;
}

int vct_queue_model(vct_ctx ctx, uint32_t model_id) {
  CTX_CHECK(ctx);
  if (model_id >= ctx->models.size()) return VCT_OK;  // ignored, renderer.cpp:211-212
  ctx->queue.push_back(model_id);
  return VCT_OK;
}

int vct_queue_point_light(vct_ctx ctx, const vct_light* light) {
  CTX_CHECK(ctx);
  if (!light) return vct_set_error(VCT_ERR_INVALID_ARG, "null light");
  if (light->kind != VCT_LIGHT_POINT && light->kind != VCT_LIGHT_DIRECTIONAL)
    return vct_set_error(VCT_ERR_INVALID_ARG, "bad light kind");
  ctx->lights.push_back(*light);
  return VCT_OK;
}

int vct_clear_queues(vct_ctx ctx) {
  CTX_CHECK(ctx);
  ctx->queue.clear();
  ctx->lights.clear();
  return VCT_OK;
}

int vct_set_rendering_phases(vct_ctx ctx, int direct, int diffuse, int specular, int shadow) {
  CTX_CHECK(ctx);
  ctx->en_direct = direct != 0;
  ctx->en_diffuse = diffuse != 0;
  ctx->en_specular = specular != 0;
  ctx->en_shadow = shadow != 0;
  return VCT_OK;
}

int vct_set_voxel_view_dir(vct_ctx ctx, int dir, float lod) {
  CTX_CHECK(ctx);
  ctx->view_voxel_dir = dir;
  ctx->view_voxel_lod = lod;
  return VCT_OK;
}

int vct_set_cone_set(vct_ctx ctx, int cone_set) {
  CTX_CHECK(ctx);
  if (cone_set < VCT_CONES_REF9 || cone_set > VCT_CONES_16) return vct_set_error(VCT_ERR_INVALID_ARG, "bad cone set");
  ctx->cone_set = cone_set;
  return VCT_OK;
}

static int timed(vct_context* ctx, int slot, int (*fn)(vct_context*)) {
  int rc = vct_pass_event(ctx, slot, false, ctx->stream);
  if (rc || (rc = fn(ctx))) return rc;
  return vct_pass_event(ctx, slot, true, ctx->stream);
}

int vct_voxelize(vct_ctx ctx) {
  CTX_CHECK(ctx);
  int rc = sync_materials(ctx);
  if (rc) return rc;
  ++ctx->volume_version;
  return timed(ctx, 0, vct_launch_voxelize);
}

int vct_mip_aniso(vct_ctx ctx) {
  CTX_CHECK(ctx);
  ++ctx->volume_version;
  return timed(ctx, 1, vct_launch_mip);
}

int vct_gbuffer(vct_ctx ctx) {
  CTX_CHECK(ctx);
  int rc = sync_materials(ctx);
  if (rc) return rc;
  return timed(ctx, 2, vct_launch_gbuffer);
}

int vct_trace(vct_ctx ctx) {
  CTX_CHECK(ctx);
  int rc = sync_materials(ctx);
  if (rc) return rc;
  return timed(ctx, 3, vct_launch_trace);
}

int vct_render(vct_ctx ctx) {
  CTX_CHECK(ctx);
  int rc;
  if ((rc = vct_frame_front(ctx))) return rc;
  if ((rc = vct_trace(ctx))) return rc;
  return vct_clear_queues(ctx);
}

size_t vct_volume_level_bytes(vct_ctx ctx, int level) {
  if (!ctx || level < 0 || level >= ctx->levels) return 0;
  uint32_t s = ctx->res >> level;
  return (size_t)s * s * s * vct::dtype_bytes(ctx->dtype);
}

// Linear (x fastest) <-> brick layout conversion on the host.
static void to_linear(const char* brick, char* lin, uint32_t s, uint32_t bpe) {
  for (uint32_t z = 0; z < s; ++z)
    for (uint32_t y = 0; y < s; ++y)
      for (uint32_t x = 0; x < s; ++x)
        memcpy(lin + (((size_t)z * s + y) * s + x) * bpe, brick + (size_t)vct::brick_index(x, y, z) * bpe, bpe);
}
static void to_brick(const char* lin, char* brick, uint32_t s, uint32_t bpe) {
  for (uint32_t z = 0; z < s; ++z)
    for (uint32_t y = 0; y < s; ++y)
      for (uint32_t x = 0; x < s; ++x)
        memcpy(brick + (size_t)vct::brick_index(x, y, z) * bpe, lin + (((size_t)z * s + y) * s + x) * bpe, bpe);
}

int vct_volume_download(vct_ctx ctx, int level, int dir, void* host, size_t bytes) {
  CTX_CHECK(ctx);
  if (level < 0 || level >= ctx->levels || dir < 0 || dir > 5 || !host)
    return vct_set_error(VCT_ERR_INVALID_ARG, "bad level/dir/pointer");
  if (bytes != vct_volume_level_bytes(ctx, level)) return vct_set_error(VCT_ERR_INVALID_ARG, "size mismatch");
  uint32_t s = ctx->res >> level;
  uint32_t bpe = vct::dtype_bytes(ctx->dtype);
  size_t dev_bytes = vct::level_elems(s) * bpe;
  std::vector<char> tmp(dev_bytes);
  VCT_HIP(hipMemcpyAsync(tmp.data(), (char*)ctx->volume.p + ctx->lvl_off[level][level == 0 ? 0 : dir], dev_bytes,
                         hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  to_linear(tmp.data(), (char*)host, s, bpe);
  return VCT_OK;
}

int vct_volume_upload(vct_ctx ctx, int level, int dir, const void* host, size_t bytes) {
  CTX_CHECK(ctx);
  if (level < 0 || level >= ctx->levels || dir < 0 || dir > 5 || !host)
    return vct_set_error(VCT_ERR_INVALID_ARG, "bad level/dir/pointer");
  if (bytes != vct_volume_level_bytes(ctx, level)) return vct_set_error(VCT_ERR_INVALID_ARG, "size mismatch");
  uint32_t s = ctx->res >> level;
  uint32_t bpe = vct::dtype_bytes(ctx->dtype);
  size_t dev_bytes = vct::level_elems(s) * bpe;
  std::vector<char> tmp(dev_bytes, 0);
  to_brick((const char*)host, tmp.data(), s, bpe);
  ++ctx->volume_version;
  VCT_HIP(hipMemcpyAsync((char*)ctx->volume.p + ctx->lvl_off[level][level == 0 ? 0 : dir], tmp.data(), dev_bytes,
                         hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_gbuffer_download(vct_ctx ctx, float* pos_depth, float* normal_mat, size_t pixels) {
  CTX_CHECK(ctx);
  if (pixels != (size_t)ctx->width * ctx->height) return vct_set_error(VCT_ERR_INVALID_ARG, "pixel count mismatch");
  if (!ctx->g_pos.p) return vct_set_error(VCT_ERR_STATE, "no G-buffer yet");
  if (pos_depth)
    VCT_HIP(hipMemcpyAsync(pos_depth, ctx->g_pos.p, pixels * 16, hipMemcpyDeviceToHost, ctx->stream));
  if (normal_mat)
    VCT_HIP(hipMemcpyAsync(normal_mat, ctx->g_nrm.p, pixels * 16, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_gbuffer_upload(vct_ctx ctx, const float* pos_depth, const float* normal_mat, size_t pixels) {
  CTX_CHECK(ctx);
  if (pixels != (size_t)ctx->width * ctx->height) return vct_set_error(VCT_ERR_INVALID_ARG, "pixel count mismatch");
  if (!pos_depth || !normal_mat) return vct_set_error(VCT_ERR_INVALID_ARG, "null pointer");
  int rc;
  if ((rc = vct_ensure(ctx->g_pos, pixels * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->g_nrm, pixels * 16, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->g_pos.p, pos_depth, pixels * 16, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->g_nrm.p, normal_mat, pixels * 16, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->g_frame.valid = false;  // no triangles behind this surface buffer
  return VCT_OK;
}

int vct_image_download(vct_ctx ctx, float* rgba, size_t pixels) {
  CTX_CHECK(ctx);
  if (!rgba) return vct_set_error(VCT_ERR_INVALID_ARG, "null pointer");
  if (pixels != (size_t)ctx->width * ctx->height) return vct_set_error(VCT_ERR_INVALID_ARG, "pixel count mismatch");
  if (!ctx->image.p) return vct_set_error(VCT_ERR_STATE, "no image yet");
  VCT_HIP(hipMemcpyAsync(rgba, ctx->image.p, pixels * 16, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_set_pass_timing(vct_ctx ctx, uint32_t mask) {
  CTX_CHECK(ctx);
  if (mask > 0x3fu) return vct_set_error(VCT_ERR_INVALID_ARG, "pass timing mask has bits past pass 5");
  ctx->time_mask = mask;
  return VCT_OK;
}

int vct_get_pass_times(vct_ctx ctx, float* ms, int n) {
  CTX_CHECK(ctx);
  if (!ms || n < 0) return vct_set_error(VCT_ERR_INVALID_ARG, "bad output");
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < 6; ++i) {
    if (ctx->ev_recorded[i]) {
      float t = 0;
      if (hipEventElapsedTime(&t, ctx->ev[2 * i], ctx->ev[2 * i + 1]) == hipSuccess) ctx->pass_ms[i] = t;
      else (void)hipGetLastError();
    }
  }
  for (int i = 0; i < n && i < 6; ++i) ms[i] = ctx->pass_ms[i];
  return VCT_OK;
}

int vct_get_voxelize_stats(vct_ctx ctx, uint64_t* fragments, uint64_t* candidates) {
  CTX_CHECK(ctx);
  if (ctx->vox_pairs_on_device) {  // a list sized without a host wait: its total is on the device
    uint32_t total = 0;
    VCT_HIP(hipMemcpyAsync(&total, (const uint32_t*)ctx->counters.p + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->last_candidates = total;
    ctx->vox_pairs_on_device = false;
  }
  if (fragments) *fragments = ctx->last_fragments;
  if (candidates) *candidates = ctx->last_candidates;
  return VCT_OK;
}

}  // extern "C"
