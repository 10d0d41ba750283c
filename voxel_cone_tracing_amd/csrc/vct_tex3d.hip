// vct_tex3d.hip - generic 3D textures behind the Texture3D entry points
// (texture_3d.h:6-10, texture_3d.cpp:3-51).
//
// A texture is one device allocation holding `levels` levels of a w x h x d
// texture in x-fastest linear order, elements of the chosen storage type.
// clear_tex_3d writes level 0 only (glClearTexImage(tex, 0, ...),
// texture_3d.cpp:42); mip_tex_3d is glGenerateMipmap (texture_3d.cpp:49):
// each level is the 2x2x2 box average of the one below (edges of odd size
// clamp to the last texel).  The radiance volume of the cone tracer is not a
// Texture3D: it has its own brick-ordered layout (vct_internal.h).
#include <cstring>

#include "vct_internal.h"

namespace vct {
namespace {

__device__ __forceinline__ float4 tload(const void* p, size_t i, int dt) {
  if (dt == VCT_RGBA8) {
    uint32_t v = reinterpret_cast<const uint32_t*>(p)[i];
    return make_float4(unorm8_exact(v & 255u), unorm8_exact((v >> 8) & 255u), unorm8_exact((v >> 16) & 255u),
                       unorm8_exact(v >> 24));
  }
  if (dt == VCT_FP16) {
    uint2 v = reinterpret_cast<const uint2*>(p)[i];
    return make_float4(f16_to_f32(v.x & 0xffffu), f16_to_f32(v.x >> 16), f16_to_f32(v.y & 0xffffu),
                       f16_to_f32(v.y >> 16));
  }
  if (dt == VCT_BF16) {
    uint2 v = reinterpret_cast<const uint2*>(p)[i];
    return make_float4(bf16_to_f32(v.x & 0xffffu), bf16_to_f32(v.x >> 16), bf16_to_f32(v.y & 0xffffu),
                       bf16_to_f32(v.y >> 16));
  }
  return reinterpret_cast<const float4*>(p)[i];
}

__device__ __forceinline__ void tstore(void* p, size_t i, int dt, float4 f) {
  if (dt == VCT_RGBA8) {
    reinterpret_cast<uint32_t*>(p)[i] =
        to_unorm8(f.x) | (to_unorm8(f.y) << 8) | (to_unorm8(f.z) << 16) | (to_unorm8(f.w) << 24);
  } else if (dt == VCT_FP16) {
    reinterpret_cast<uint2*>(p)[i] = make_uint2((uint32_t)f32_to_f16(f.x) | ((uint32_t)f32_to_f16(f.y) << 16),
                                                (uint32_t)f32_to_f16(f.z) | ((uint32_t)f32_to_f16(f.w) << 16));
  } else if (dt == VCT_BF16) {
    reinterpret_cast<uint2*>(p)[i] = make_uint2((uint32_t)f32_to_bf16(f.x) | ((uint32_t)f32_to_bf16(f.y) << 16),
                                                (uint32_t)f32_to_bf16(f.z) | ((uint32_t)f32_to_bf16(f.w) << 16));
  } else {
    reinterpret_cast<float4*>(p)[i] = f;
  }
}

__global__ void tex_fill_kernel(void* p, size_t n, int dt, float4 v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    tstore(p, i, dt, v);
}

__global__ void tex_box_kernel(const void* src, int sw, int sh, int sd, void* dst, int dw, int dh, int dd, int dt) {
#pragma clang fp contract(off)
  const size_t n = (size_t)dw * dh * dd;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % dw), y = (int)((i / dw) % dh), z = (int)(i / ((size_t)dw * dh));
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < 8; ++k) {
      const int sx = min(2 * x + (k & 1), sw - 1), sy = min(2 * y + ((k >> 1) & 1), sh - 1),
                sz = min(2 * z + (k >> 2), sd - 1);
      const float4 v = tload(src, ((size_t)sz * sh + sy) * sw + sx, dt);
      acc = make_float4(acc.x + v.x, acc.y + v.y, acc.z + v.z, acc.w + v.w);
    }
    tstore(dst, i, dt, make_float4(acc.x * 0.125f, acc.y * 0.125f, acc.z * 0.125f, acc.w * 0.125f));
  }
}

}  // namespace
}  // namespace vct

using namespace vct;

int vct_tex3d_clear_launch(vct_context* ctx, VctTex3D& t, const float rgba[4]) {
  const size_t n = (size_t)t.w * t.h * t.d;
  hipLaunchKernelGGL(tex_fill_kernel, dim3(1024), dim3(256), 0, ctx->stream, t.dptr, n, t.dtype,
                     make_float4(rgba[0], rgba[1], rgba[2], rgba[3]));
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

int vct_tex3d_mip_launch(vct_context* ctx, VctTex3D& t) {
  for (int l = 1; l < t.levels; ++l) {
    hipLaunchKernelGGL(tex_box_kernel, dim3(1024), dim3(256), 0, ctx->stream, (const char*)t.dptr + t.off[l - 1],
                       t.lw[l - 1], t.lh[l - 1], t.ld[l - 1], (char*)t.dptr + t.off[l], t.lw[l], t.lh[l], t.ld[l],
                       t.dtype);
  }
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

static VctTex3D* get_tex(vct_context* ctx, uint32_t tex) {
  if (tex == 0 || tex > ctx->textures.size()) return nullptr;
  VctTex3D& t = ctx->textures[tex - 1];
  return t.alive ? &t : nullptr;
}

extern "C" {

int vct_tex3d_create(vct_ctx ctx, int w, int h, int d, int levels, int dtype, uint32_t* out) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  *out = 0;
  if (w < 1 || h < 1 || d < 1 || w > 2048 || h > 2048 || d > 2048)
    return vct_set_error(VCT_ERR_INVALID_ARG, "texture size must be in [1, 2048]");
  if (dtype < VCT_RGBA8 || dtype > VCT_FP32) return vct_set_error(VCT_ERR_INVALID_ARG, "bad dtype");
  int maxl = 1;
  for (int m = w > h ? (w > d ? w : d) : (h > d ? h : d); m > 1; m >>= 1) ++maxl;
  if (levels < 1 || levels > maxl) return vct_set_error(VCT_ERR_INVALID_ARG, "bad level count");
  VCT_HIP(hipSetDevice(ctx->device));
  VctTex3D t;
  t.w = w;
  t.h = h;
  t.d = d;
  t.levels = levels;
  t.dtype = dtype;
  size_t off = 0;
  int cw = w, ch = h, cd = d;
  for (int l = 0; l < levels; ++l) {
    t.off.push_back(off);
    t.lw.push_back(cw);
    t.lh.push_back(ch);
    t.ld.push_back(cd);
    off += (((size_t)cw * ch * cd * dtype_bytes(dtype)) + 255) & ~(size_t)255;
    cw = cw > 1 ? cw / 2 : 1;
    ch = ch > 1 ? ch / 2 : 1;
    cd = cd > 1 ? cd / 2 : 1;
  }
  t.bytes = off;
  hipError_t e = hipMalloc(&t.dptr, off);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return vct_set_error(VCT_ERR_OUT_OF_MEMORY, "texture allocation failed");
  }
  VCT_HIP(hipMemsetAsync(t.dptr, 0, off, ctx->stream));
  t.alive = true;
  ctx->textures.push_back(std::move(t));
  *out = (uint32_t)ctx->textures.size();  // ids start at 1, like GL names
  return VCT_OK;
}

int vct_tex3d_destroy(vct_ctx ctx, uint32_t tex) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  VctTex3D* t = get_tex(ctx, tex);
  if (!t) return VCT_OK;  // glDeleteTextures ignores unknown names
  VCT_HIP(hipSetDevice(ctx->device));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  VCT_HIP(hipFree(t->dptr));
  t->dptr = nullptr;
  t->alive = false;
  return VCT_OK;
}

int vct_tex3d_activate(vct_ctx ctx, uint32_t tex, uint32_t unit) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  if (tex != 0 && !get_tex(ctx, tex)) return vct_set_error(VCT_ERR_BAD_HANDLE, "unknown texture");
  if (unit >= 32) return vct_set_error(VCT_ERR_INVALID_ARG, "texture unit out of range");
  return VCT_OK;
}

int vct_tex3d_clear(vct_ctx ctx, uint32_t tex, const float rgba[4]) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  VctTex3D* t = get_tex(ctx, tex);
  if (!t) return vct_set_error(VCT_ERR_BAD_HANDLE, "unknown texture");
  if (!rgba) return vct_set_error(VCT_ERR_INVALID_ARG, "null colour");
  VCT_HIP(hipSetDevice(ctx->device));
  return vct_tex3d_clear_launch(ctx, *t, rgba);
}

int vct_tex3d_mip(vct_ctx ctx, uint32_t tex) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  VctTex3D* t = get_tex(ctx, tex);
  if (!t) return vct_set_error(VCT_ERR_BAD_HANDLE, "unknown texture");
  VCT_HIP(hipSetDevice(ctx->device));
  return vct_tex3d_mip_launch(ctx, *t);
}

int vct_tex3d_upload(vct_ctx ctx, uint32_t tex, int level, const void* host, size_t bytes) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  VctTex3D* t = get_tex(ctx, tex);
  if (!t) return vct_set_error(VCT_ERR_BAD_HANDLE, "unknown texture");
  if (level < 0 || level >= t->levels || !host) return vct_set_error(VCT_ERR_INVALID_ARG, "bad level or pointer");
  size_t n = (size_t)t->lw[level] * t->lh[level] * t->ld[level] * dtype_bytes(t->dtype);
  if (bytes != n) return vct_set_error(VCT_ERR_INVALID_ARG, "size mismatch");
  VCT_HIP(hipSetDevice(ctx->device));
  VCT_HIP(hipMemcpyAsync((char*)t->dptr + t->off[level], host, n, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_tex3d_download(vct_ctx ctx, uint32_t tex, int level, void* host, size_t bytes) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  VctTex3D* t = get_tex(ctx, tex);
  if (!t) return vct_set_error(VCT_ERR_BAD_HANDLE, "unknown texture");
  if (level < 0 || level >= t->levels || !host) return vct_set_error(VCT_ERR_INVALID_ARG, "bad level or pointer");
  size_t n = (size_t)t->lw[level] * t->lh[level] * t->ld[level] * dtype_bytes(t->dtype);
  if (bytes != n) return vct_set_error(VCT_ERR_INVALID_ARG, "size mismatch");
  VCT_HIP(hipSetDevice(ctx->device));
  VCT_HIP(hipMemcpyAsync(host, (char*)t->dptr + t->off[level], n, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

}  // extern "C"
