// vct_mip.hip - anisotropic mip-chain builder (kernel b).
//
// Computes the same values as Renderer::filter (renderer.cpp:283-314)
// dispatching mipmap.comp (mipmap.comp:45-99) once per level: every texel of
// level k+1 in direction map d is the mean of four front-to-back alpha blends
// alpha_blend(f, b) = f + (1 - f.a) * b (mipmap.comp:40-43) of the 2x2x2 block
// below it, paired along the axis of d, then converted to the storage type.
//
// MI355X design: one workgroup reads one aligned 16^3 block of its source
// level (64 consecutive bricks, a single contiguous range, coalesced 16-byte
// loads) into LDS, then produces up to four levels of all six maps from LDS
// (8^3, 4^3, 2^3, 1 texels per map), writing each output brick range once.
// Each level is computed from the stored (rounded) values of the level
// below, exactly as the reference's per-level dispatches do, so the fused
// result equals the level-by-level one bit for bit.  Level 0 is read once
// from HBM; every mip texel is written once.  A second launch continues from
// the level the first one stopped at when more levels are requested.
#include "vct_internal.h"

namespace vct {
namespace {

// Pair tables of mipmap.comp:58-98.  Source texel i of the 2x2x2 block sits
// at block + voxel_offsets[i] (mipmap.comp:10-20): i = 4*(1-x) + 2*(1-y) + (1-z).
// For each direction: four (front, back) pairs, summed in this order.
constexpr uint8_t kPairs[6][4][2] = {
    {{0, 4}, {1, 5}, {2, 6}, {3, 7}},  // -x
    {{4, 0}, {5, 1}, {6, 2}, {7, 3}},  // +x
    {{0, 2}, {1, 3}, {5, 7}, {4, 6}},  // -y
    {{2, 0}, {3, 1}, {7, 5}, {6, 4}},  // +y
    {{0, 1}, {2, 3}, {4, 5}, {6, 7}},  // -z
    {{1, 0}, {3, 2}, {5, 4}, {7, 6}},  // +z
};

template <int DT>
struct Codec;

template <>
struct Codec<VCT_RGBA8> {
  using T = uint32_t;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(unorm8_exact(v & 255u), unorm8_exact((v >> 8) & 255u),
                       unorm8_exact((v >> 16) & 255u), unorm8_exact(v >> 24));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return to_unorm8(f.x) | (to_unorm8(f.y) << 8) | (to_unorm8(f.z) << 16) | (to_unorm8(f.w) << 24);
  }
};
template <>
struct Codec<VCT_FP16> {
  using T = uint2;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(f16_to_f32(v.x & 0xffffu), f16_to_f32(v.x >> 16), f16_to_f32(v.y & 0xffffu),
                       f16_to_f32(v.y >> 16));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return make_uint2((uint32_t)f32_to_f16(f.x) | ((uint32_t)f32_to_f16(f.y) << 16),
                      (uint32_t)f32_to_f16(f.z) | ((uint32_t)f32_to_f16(f.w) << 16));
  }
};
template <>
struct Codec<VCT_BF16> {
  using T = uint2;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(bf16_to_f32(v.x & 0xffffu), bf16_to_f32(v.x >> 16), bf16_to_f32(v.y & 0xffffu),
                       bf16_to_f32(v.y >> 16));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return make_uint2((uint32_t)f32_to_bf16(f.x) | ((uint32_t)f32_to_bf16(f.y) << 16),
                      (uint32_t)f32_to_bf16(f.z) | ((uint32_t)f32_to_bf16(f.w) << 16));
  }
};
template <>
struct Codec<VCT_FP32> {
  using T = float4;
  static __device__ __forceinline__ float4 dec(T v) { return v; }
  static __device__ __forceinline__ T enc(float4 f) { return f; }
};

// f + (1 - f.a) * b, each product and sum rounded separately (no fused
// multiply-add), as the oracle evaluates it.
__device__ __forceinline__ float4 alpha_blend(float4 f, float4 b) {
#pragma clang fp contract(off)
  float k = 1.0f - f.w;
  return make_float4(f.x + k * b.x, f.y + k * b.y, f.z + k * b.z, f.w + k * b.w);
}
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
#pragma clang fp contract(off)
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// One mip texel of direction D from the eight source texels:
// ((b0 + b1) + b2) + b3, then / 4 (mipmap.comp:59-98).
template <int D>
__device__ __forceinline__ float4 dir_texel(const float4 (&v)[8]) {
  float4 acc = alpha_blend(v[kPairs[D][0][0]], v[kPairs[D][0][1]]);
  acc = add4(acc, alpha_blend(v[kPairs[D][1][0]], v[kPairs[D][1][1]]));
  acc = add4(acc, alpha_blend(v[kPairs[D][2][0]], v[kPairs[D][2][1]]));
  acc = add4(acc, alpha_blend(v[kPairs[D][3][0]], v[kPairs[D][3][1]]));
  return make_float4(acc.x / 4.0f, acc.y / 4.0f, acc.z / 4.0f, acc.w / 4.0f);
}

__device__ __forceinline__ float4 dir_texel_rt(int d, const float4 (&v)[8]) {
  switch (d) {
    case 0: return dir_texel<0>(v);
    case 1: return dir_texel<1>(v);
    case 2: return dir_texel<2>(v);
    case 3: return dir_texel<3>(v);
    case 4: return dir_texel<4>(v);
    default: return dir_texel<5>(v);
  }
}

struct MipArgs {
  const void* src[6];   // source level per direction (all equal when shared)
  void* dst[4][6];      // up to four output levels per direction
  uint32_t src_size;    // source level edge length
  uint32_t block;       // edge of the source block one workgroup handles (<= 16)
  int nout;             // output levels produced by this launch (1..4)
  int shared_src;       // 1: one source for all six maps (level 0)
};

// Index of local coordinate (x,y,z) inside a block of edge n stored like a
// level: brick order when n >= 4, x-fastest inside one padded brick otherwise.
__device__ __forceinline__ uint32_t local_index(uint32_t n, uint32_t x, uint32_t y, uint32_t z) {
  return brick_index(x, y, z);  // same formula covers the padded case (x,y,z < 4)
  (void)n;
}

template <int DT>
__global__ __launch_bounds__(256) void mip_fused_kernel(MipArgs a) {
  using C = Codec<DT>;
  using T = typename C::T;
  // LDS: source block (4096 texels) + per direction level outputs.
  // 8^3 + 4^3 + 2^3 + 1 = 585 texels per direction for the following levels;
  // each output level is padded to at least one brick (64) for indexing.
  constexpr int kSrc = 4096;
  constexpr int kL1 = 512, kL2 = 64, kL3 = 64, kL4 = 64;
  constexpr int kPerDir = kL1 + kL2 + kL3 + kL4;
  __shared__ T s_src[kSrc];
  __shared__ T s_out[6 * kPerDir];
  const int lvl_base[4] = {0, kL1, kL1 + kL2, kL1 + kL2 + kL3};

  const uint32_t B = a.block;
  const uint32_t nblk_edge = a.src_size / B;
  const uint32_t blk = blockIdx.x;
  const int dir_lo = a.shared_src ? 0 : (int)blockIdx.y;
  const int dir_hi = a.shared_src ? 6 : (int)blockIdx.y + 1;
  const int src_dir = dir_lo;

  // Block origin: blocks are enumerated in Morton order so that block `blk`
  // covers source elements [blk * B^3, (blk+1) * B^3) when B >= 4.
  uint32_t bx, by, bz;
  {
    uint32_t m = blk;
    bx = compact3(m);
    by = compact3(m >> 1);
    bz = compact3(m >> 2);
  }
  (void)nblk_edge;

  // 1) stage the source block
  const T* src = reinterpret_cast<const T*>(a.src[src_dir]);
  const uint32_t n_src = B >= 4 ? B * B * B : 64u;
  const uint64_t src_base = B >= 4 ? (uint64_t)blk * n_src : 0ull;
  for (uint32_t i = threadIdx.x; i < n_src; i += blockDim.x) s_src[i] = src[src_base + i];
  __syncthreads();

  // 2) levels
  uint32_t n_in = B;
  for (int l = 0; l < a.nout; ++l) {
    const uint32_t n_out = n_in >> 1;
    const uint32_t cnt = n_out * n_out * n_out;
    const uint32_t out_size = (a.src_size >> (l + 1));
    const uint32_t ox0 = bx * n_out, oy0 = by * n_out, oz0 = bz * n_out;
    for (uint32_t p = threadIdx.x; p < cnt; p += blockDim.x) {
      uint32_t x, y, z;
      if (n_out >= 4) {
        brick_coords(p, x, y, z);
      } else {
        x = p % n_out;
        y = (p / n_out) % n_out;
        z = p / (n_out * n_out);
      }
      const uint32_t gi = brick_index(ox0 + x, oy0 + y, oz0 + z);
      if (a.shared_src && l == 0) {
        // all six maps from the same eight texels
        float4 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const uint32_t sx = 2 * x + ((i & 4) ? 0 : 1);
          const uint32_t sy = 2 * y + ((i & 2) ? 0 : 1);
          const uint32_t sz = 2 * z + ((i & 1) ? 0 : 1);
          v[i] = C::dec(s_src[local_index(n_in, sx, sy, sz)]);
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          const T e = C::enc(dir_texel_rt(d, v));
          s_out[d * kPerDir + lvl_base[l] + local_index(n_out, x, y, z)] = e;
          reinterpret_cast<T*>(a.dst[l][d])[gi] = e;
        }
      } else {
        for (int d = dir_lo; d < dir_hi; ++d) {
          float4 v[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint32_t sx = 2 * x + ((i & 4) ? 0 : 1);
            const uint32_t sy = 2 * y + ((i & 2) ? 0 : 1);
            const uint32_t sz = 2 * z + ((i & 1) ? 0 : 1);
            const uint32_t li = local_index(n_in, sx, sy, sz);
            const T raw = (l == 0) ? s_src[li] : s_out[d * kPerDir + lvl_base[l - 1] + li];
            v[i] = C::dec(raw);
          }
          const T e = C::enc(dir_texel_rt(d, v));
          s_out[d * kPerDir + lvl_base[l] + local_index(n_out, x, y, z)] = e;
          reinterpret_cast<T*>(a.dst[l][d])[gi] = e;
        }
      }
    }
    (void)out_size;
    __syncthreads();
    n_in = n_out;
  }
}

template <int DT>
int launch_chain(vct_context* ctx) {
  const uint32_t res = ctx->res;
  const int L = ctx->levels;
  int src_level = 0;
  char* base = reinterpret_cast<char*>(ctx->volume.p);
  while (src_level < L - 1) {
    const uint32_t s = res >> src_level;
    const uint32_t B = s < 16 ? s : 16;
    int nout = 0;
    for (uint32_t b = B; b > 1 && nout < 4 && src_level + nout < L - 1; b >>= 1) ++nout;
    if (nout == 0) break;
    MipArgs a;
    for (int d = 0; d < 6; ++d) a.src[d] = base + ctx->lvl_off[src_level][src_level == 0 ? 0 : d];
    for (int l = 0; l < 4; ++l)
      for (int d = 0; d < 6; ++d)
        a.dst[l][d] = (l < nout) ? (void*)(base + ctx->lvl_off[src_level + 1 + l][d]) : nullptr;
    a.src_size = s;
    a.block = B;
    a.nout = nout;
    a.shared_src = src_level == 0 ? 1 : 0;
    const uint32_t nb = (s / B) * (s / B) * (s / B);
    dim3 grid(nb, a.shared_src ? 1 : 6);
    hipLaunchKernelGGL(mip_fused_kernel<DT>, grid, dim3(256), 0, ctx->stream, a);
    VCT_HIP(hipGetLastError());
    src_level += nout;
  }
  return VCT_OK;
}

}  // namespace
}  // namespace vct

int vct_launch_mip(vct_context* ctx) {
  switch (ctx->dtype) {
    case VCT_RGBA8: return vct::launch_chain<VCT_RGBA8>(ctx);
    case VCT_FP16: return vct::launch_chain<VCT_FP16>(ctx);
    case VCT_BF16: return vct::launch_chain<VCT_BF16>(ctx);
    case VCT_FP32: return vct::launch_chain<VCT_FP32>(ctx);
  }
  return vct_set_error(VCT_ERR_INVALID_ARG, "bad dtype");
}
