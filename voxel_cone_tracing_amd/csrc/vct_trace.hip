// vct_trace.hip - per-pixel cone marcher (kernel c).
//
// Shades every covered pixel as voxel_cone_tracing.frag:246-275 does:
//   pos = scale_and_bias(world / cube_size)                      (:248)
//   diffuse  = Kd * trace_diffuse(pos, n)                         (:140-168)
//   direct   = direct_light(pos, view_dir) with shadow cones       (:175-218)
//   specular = Ks * trace_specular(pos, n, view_dir)              (:229-235)
//   refraction override for illum 4/6/7/9                         (:273-274)
// trace_cone (:88-119) marches front to back, c += (1 - c.a) * s, with
// mip = max(log2(diam * res), 0), step = max(diam / 2, voxel_size), and
// sample_voxel (:80-86) = |d.x| T[x] + |d.y| T[y] + |d.z| T[z] (unnormalised).
//
// Sampler: a software quadrilinear textureLod over the brick/Morton volume
// with GL semantics (texture_3d.cpp:10-15): texel centres at (i + 0.5) / s,
// CLAMP_TO_BORDER with a zero border, LINEAR_MIPMAP_LINEAR between floor(lod)
// and floor(lod) + 1, lod clamped to [0, L-1].  The three maps of one sample
// share coordinates and weights, so corner addresses are computed once per
// level and all 24 corner loads of a level are issued before the first is
// used; level 0 is shared by all maps, so it is fetched once per step.
//
// Work mapping: a workgroup of 4 waves shades an 8x8 pixel block (lane =
// pixel); the cones of a pixel are dealt over the 4 waves, and all lanes of
// a wave march the same cone at the same time over neighbouring surface
// points, so their texel fetches fall on the same bricks and L2 lines.
// Accumulation is FP32 for every storage type.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "vct_internal.h"

namespace vct {

constexpr int kMaxCones = 16;

struct ConeSet {
  float a[kMaxCones], b[kMaxCones], c[kMaxCones];  // dir = a n + b o1 + c o2
  float aperture;
  int n;
};

struct TraceArgs {
  const char* vol;
  uint64_t off[16][6];
  int levels;
  int res;
  const float4* gpos;
  const float4* gnrm;
  const DevMaterial* mats;
  uint32_t nmat;
  float4* image;
  int W, H;
  float cube_size;
  float cam[3];
  int en_direct, en_diffuse, en_specular, en_shadow;
  int view_dir;
  float view_lod;
  int demodulate;  // 1: write final / max(Kd, eps) (the learned path's target)
  unsigned long long* step_counter;  // optional: cone steps taken (diagnostics)
  // pixel selection (both null: every pixel, image indexed like the G-buffer)
  const uint32_t* pix_list;    // listed pixels (y * W + x); image[k] receives pixel k
  uint32_t npix_list;          // count (host), or the bound of *pix_count when that is set
  const uint32_t* pix_count;   // optional device-side count of listed pixels
  int pix_scatter;             // 1: image[pixel] receives the pixel instead of image[k]
  const uint32_t* tile_list;   // listed 32x32 tiles (ty * tiles_x + tx); 16 workgroups per tile
  const uint32_t* tile_count;  // device-side number of listed tiles
  int tiles_x;
  const uint32_t* tile_flag;   // optional: listed pixels are traced only where tile_flag[tile] == flag_want
  uint32_t flag_want;
  int grid_stride;             // tile-list mode: 2 = the odd-odd pixels of the tiles (a 16x16 grid per tile),
                               // written to image[(y >> 1) * half_w + (x >> 1)]
  int half_w;
  ConeSet cones;
  LightSet lights;
};

// Storage types: raw texel as loaded (Raw) and its decode to float4; the
// decode folds into the blend (v_fma_mix_f32 for the 16-bit types).
template <int DT>
struct Tex;
template <>
struct Tex<VCT_RGBA8> {
  using Raw = uint32_t;
  static __device__ __forceinline__ float4 decode(Raw v) {
    const float k = 1.0f / 255.0f;
    return make_float4((float)(v & 255u) * k, (float)((v >> 8) & 255u) * k, (float)((v >> 16) & 255u) * k,
                       (float)(v >> 24) * k);
  }
};
template <>
struct Tex<VCT_FP16> {
  using Raw = uint2;
  static __device__ __forceinline__ float4 decode(Raw v) {
    return make_float4(f16_to_f32(v.x & 0xffffu), f16_to_f32(v.x >> 16), f16_to_f32(v.y & 0xffffu),
                       f16_to_f32(v.y >> 16));
  }
};
template <>
struct Tex<VCT_BF16> {
  using Raw = uint2;
  static __device__ __forceinline__ float4 decode(Raw v) {
    return make_float4(bf16_to_f32(v.x & 0xffffu), bf16_to_f32(v.x >> 16), bf16_to_f32(v.y & 0xffffu),
                       bf16_to_f32(v.y >> 16));
  }
};
template <>
struct Tex<VCT_FP32> {
  using Raw = float4;
  static __device__ __forceinline__ float4 decode(Raw v) { return v; }
};

// x-paired copy of the volume read by the marcher: entry i of a map holds
// texel i and its +x neighbour (zero past the last column), so one load
// returns both x-corners of a trilinear fetch - half the load instructions,
// which the texture address units otherwise bound.  16-bit storage pairs
// channel by channel (four words of (c(x), c(x+1))), RGBA8 as two words;
// FP32 texels are already 16 bytes and are not paired.
template <int DT>
struct Pair {
  static constexpr bool kOn = false;
};
template <>
struct Pair<VCT_RGBA8> {
  static constexpr bool kOn = true;
  using Raw = uint2;
  static __device__ __forceinline__ Raw make(uint32_t a, uint32_t b) { return pair_pack(a, b); }
  static __device__ __forceinline__ float4 first(Raw v) { return Tex<VCT_RGBA8>::decode(v.x); }
  static __device__ __forceinline__ float4 second(Raw v) { return Tex<VCT_RGBA8>::decode(v.y); }
};
template <int DT>
struct Pair16 {
  static constexpr bool kOn = true;
  using Raw = uint4;
  static __device__ __forceinline__ Raw make(uint2 a, uint2 b) { return pair_pack(a, b); }
  static __device__ __forceinline__ float4 first(Raw v) {
    return Tex<DT>::decode(make_uint2((v.x & 0xffffu) | (v.y << 16), (v.z & 0xffffu) | (v.w << 16)));
  }
  static __device__ __forceinline__ float4 second(Raw v) {
    return Tex<DT>::decode(make_uint2((v.x >> 16) | (v.y & 0xffff0000u), (v.z >> 16) | (v.w & 0xffff0000u)));
  }
};
template <>
struct Pair<VCT_FP16> : Pair16<VCT_FP16> {};
template <>
struct Pair<VCT_BF16> : Pair16<VCT_BF16> {};

__device__ __forceinline__ float4 f4(float x) { return make_float4(x, x, x, x); }
__device__ __forceinline__ float4 operator+(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 operator*(float s, float4 a) {
  return make_float4(s * a.x, s * a.y, s * a.z, s * a.w);
}

// Per-workgroup sampler tables in LDS: the byte offset of a texel within a
// map, brick_index(x, y, z) << shift = sx[x] + sy[y] + sz[z] (spread3 of each
// coordinate, pre-shifted by the axis and the entry size, for edges up to
// 512: the volume's 4 GiB bound rules out 1024^3 in every storage type, and
// the launch checks it), and the byte offset of every (level, map) of the
// volume (a lane's level varies, so the table is indexed per lane; LDS reads
// are far cheaper than per-lane reads of kernel arguments).  Offsets are
// 32-bit: the volume stays below 4 GiB (checked when it is allocated), so
// every texel load is the uniform volume pointer plus one 32-bit lane offset.
constexpr int kMaxEdge = 512;
struct SamplerLds {
  uint32_t sx[kMaxEdge], sy[kMaxEdge], sz[kMaxEdge];
  uint32_t off[16][6];
};

template <int DT>
struct Fetch {  // the unit one texel load returns: a pair where the volume is paired
  using Raw = typename Tex<DT>::Raw;
};
template <>
struct Fetch<VCT_RGBA8> {
  using Raw = Pair<VCT_RGBA8>::Raw;
};
template <>
struct Fetch<VCT_FP16> {
  using Raw = Pair<VCT_FP16>::Raw;
};
template <>
struct Fetch<VCT_BF16> {
  using Raw = Pair<VCT_BF16>::Raw;
};
template <int DT>
constexpr int kTexShift = sizeof(typename Fetch<DT>::Raw) == 16 ? 4 : (sizeof(typename Fetch<DT>::Raw) == 8 ? 3 : 2);

template <int DT>
__device__ __forceinline__ void sampler_init(const TraceArgs& A, SamplerLds* S) {
  constexpr int sh = kTexShift<DT>;
  for (int q = threadIdx.x; q < kMaxEdge; q += blockDim.x) {
    const uint32_t v = spread3((uint32_t)q);
    S->sx[q] = v << sh;
    S->sy[q] = v << (sh + 1);
    S->sz[q] = v << (sh + 2);
  }
  for (int q = threadIdx.x; q < 16 * 6; q += blockDim.x) S->off[q / 6][q % 6] = (uint32_t)A.off[q / 6][q % 6];
}

// Corner byte offsets within a map: spread parts have disjoint bits, so
// brick_index(x, y, z) << shift = ix[dx] + iyz[dy + 2 dz].  Paired volumes
// use ix[0] only: the entry of the clamped x0 holds both x-corners, weighted
// w[2k] (first texel) and w[2k + 1] (second) for the (y, z) corner k.
struct Corners {
  uint32_t ix[2], iyz[4];
  float w[8];  // corner weights (dx + 2 dy + 4 dz); zero for border texels
};

template <int DT>
__device__ __forceinline__ void corners(const SamplerLds* S, float3 p, int s, Corners& c) {
  const float fs = (float)s;
  const float u = p.x * fs - 0.5f, v = p.y * fs - 0.5f, w = p.z * fs - 0.5f;
  const float fu = floorf(u), fv = floorf(v), fw = floorf(w);
  const float au = u - fu, av = v - fv, aw = w - fw;
  const int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
  // Out-of-range texels take the zero border colour: weight 0, any valid
  // index (0 <= i < s as one unsigned compare).
  const uint32_t us = (uint32_t)s;
  float wx0 = (uint32_t)x0 < us ? 1.0f - au : 0.0f;
  float wx1 = (uint32_t)(x0 + 1) < us ? au : 0.0f;
  if (Pair<DT>::kOn && x0 == -1) {  // the pair at column 0 holds x1 = 0 first
    wx0 = au;
    wx1 = 0.0f;
  }
  const float wy0 = (uint32_t)y0 < us ? 1.0f - av : 0.0f;
  const float wy1 = (uint32_t)(y0 + 1) < us ? av : 0.0f;
  const float wz0 = (uint32_t)z0 < us ? 1.0f - aw : 0.0f;
  const float wz1 = (uint32_t)(z0 + 1) < us ? aw : 0.0f;
  const float w00 = wy0 * wz0, w10 = wy1 * wz0, w01 = wy0 * wz1, w11 = wy1 * wz1;
  c.w[0] = wx0 * w00;
  c.w[1] = wx1 * w00;
  c.w[2] = wx0 * w10;
  c.w[3] = wx1 * w10;
  c.w[4] = wx0 * w01;
  c.w[5] = wx1 * w01;
  c.w[6] = wx0 * w11;
  c.w[7] = wx1 * w11;
  // Level edges are powers of two, so i & (s - 1) is a valid index for any i;
  // it is the texel itself whenever the weight is not zero.  Only x0 = -1
  // must map to column 0 (the paired entry there holds its right corner).
  const int m = s - 1;
  c.ix[0] = S->sx[max(x0, 0) & m];
  c.ix[1] = S->sx[(x0 + 1) & m];
  const uint32_t y0p = S->sy[y0 & m];
  const uint32_t y1p = S->sy[(y0 + 1) & m];
  const uint32_t z0p = S->sz[z0 & m];
  const uint32_t z1p = S->sz[(z0 + 1) & m];
  c.iyz[0] = y0p + z0p;
  c.iyz[1] = y1p + z0p;
  c.iyz[2] = y0p + z1p;
  c.iyz[3] = y1p + z1p;
}

// The eight corners of one map (byte offset `base`), all loads issued before
// any is used (border corners have weight 0 and a clamped, valid index: no
// per-corner branches).
template <int DT>
__device__ __forceinline__ void gather8(const char* __restrict__ vol, uint32_t base, const Corners& c,
                                        typename Tex<DT>::Raw (&r)[8]) {
  using Raw = typename Tex<DT>::Raw;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = *reinterpret_cast<const Raw*>(vol + (base + c.ix[k & 1] + c.iyz[k >> 1]));
}

template <int DT>
__device__ __forceinline__ void gather4(const char* __restrict__ vol, uint32_t base, const Corners& c,
                                        typename Pair<DT>::Raw (&r)[4]) {
  using Raw = typename Pair<DT>::Raw;
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = *reinterpret_cast<const Raw*>(vol + (base + c.ix[0] + c.iyz[k]));
}

// Weights of the four (y, z) corner pairs, as the packed 16-bit operand of a
// two-term dot product for 16-bit storage (rounded to nearest in the storage
// type: the filter weights carry the precision of the stored texels).
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));

template <int DT>
struct PairWeights {
  float w[8];
  __device__ __forceinline__ explicit PairWeights(const Corners& c) {
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = c.w[k];
  }
};
template <>
struct PairWeights<VCT_FP16> {
  half2_t w[4];
  __device__ __forceinline__ explicit PairWeights(const Corners& c) {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = half2_t{(_Float16)c.w[2 * k], (_Float16)c.w[2 * k + 1]};
  }
};
template <>
struct PairWeights<VCT_BF16> {
  bf162_t w[4];
  __device__ __forceinline__ explicit PairWeights(const Corners& c) {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = bf162_t{(__bf16)c.w[2 * k], (__bf16)c.w[2 * k + 1]};
  }
};

// Blend of the eight corners of one paired map: per channel, one FP32
// two-term dot product (v_dot2_f32_f16 / _bf16) per corner pair for 16-bit
// storage; RGBA8 decodes both texels and accumulates in FP32.
template <int DT>
__device__ __forceinline__ float4 blend4(const PairWeights<DT>& W, const typename Pair<DT>::Raw (&r)[4]) {
  float4 acc = f4(0.0f);
  if constexpr (DT == VCT_FP16) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc.x = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, r[k].x), W.w[k], acc.x, false);
      acc.y = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, r[k].y), W.w[k], acc.y, false);
      acc.z = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, r[k].z), W.w[k], acc.z, false);
      acc.w = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, r[k].w), W.w[k], acc.w, false);
    }
  } else if constexpr (DT == VCT_BF16) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc.x = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf162_t, r[k].x), W.w[k], acc.x, false);
      acc.y = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf162_t, r[k].y), W.w[k], acc.y, false);
      acc.z = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf162_t, r[k].z), W.w[k], acc.z, false);
      acc.w = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf162_t, r[k].w), W.w[k], acc.w, false);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      acc = acc + W.w[2 * k] * Pair<DT>::first(r[k]) + W.w[2 * k + 1] * Pair<DT>::second(r[k]);
  }
  return acc;
}

template <int DT>
__device__ __forceinline__ float4 blend8(const Corners& c, const typename Tex<DT>::Raw (&r)[8]) {
  float4 acc = f4(0.0f);
#pragma unroll
  for (int k = 0; k < 8; ++k) acc = acc + c.w[k] * Tex<DT>::decode(r[k]);
  return acc;
}

// One level of sample_voxel: level 0 is the one map shared by all directions
// (|d.x| + |d.y| + |d.z| times one trilinear fetch), a higher level blends
// the three directional maps mx, my, mz.
template <int DT>
__device__ __forceinline__ float4 sample_level(const TraceArgs& A, const SamplerLds* S, float3 p, float3 ad, int mx,
                                               int my, int mz, int l) {
  Corners c;
  corners<DT>(S, p, A.res >> l, c);
  if constexpr (Pair<DT>::kOn) {
    using PRaw = typename Pair<DT>::Raw;
    const PairWeights<DT> W(c);
    if (l == 0) {
      PRaw r[4];
      gather4<DT>(A.vol, 0u, c, r);
      return (ad.x + ad.y + ad.z) * blend4<DT>(W, r);
    }
    PRaw rx[4], ry[4], rz[4];
    gather4<DT>(A.vol, S->off[l][mx], c, rx);
    gather4<DT>(A.vol, S->off[l][my], c, ry);
    gather4<DT>(A.vol, S->off[l][mz], c, rz);
    return ad.x * blend4<DT>(W, rx) + ad.y * blend4<DT>(W, ry) + ad.z * blend4<DT>(W, rz);
  }
  using Raw = typename Tex<DT>::Raw;
  if (l == 0) {
    Raw r[8];
    gather8<DT>(A.vol, 0u, c, r);
    return (ad.x + ad.y + ad.z) * blend8<DT>(c, r);
  }
  // 16-byte FP32 texels: one map in flight at a time (24 would spill)
  Raw r[8];
  gather8<DT>(A.vol, S->off[l][mx], c, r);
  float4 acc = ad.x * blend8<DT>(c, r);
  gather8<DT>(A.vol, S->off[l][my], c, r);
  acc = acc + ad.y * blend8<DT>(c, r);
  gather8<DT>(A.vol, S->off[l][mz], c, r);
  return acc + ad.z * blend8<DT>(c, r);
}

// sample_voxel (voxel_cone_tracing.frag:80-86) with its three textureLod
// calls folded together: lod clamped to [0, L-1], levels floor(lod) and
// floor(lod) + 1 blended by the fraction (the second only when it is > 0).
template <int DT>
__device__ __forceinline__ float4 sample_voxel(const TraceArgs& A, const SamplerLds* S, float3 p, float3 ad, int mx,
                                               int my, int mz, float lod) {
  const float lmax = (float)(A.levels - 1);
  lod = fminf(fmaxf(lod, 0.0f), lmax);
  const int l0 = (int)floorf(lod);
  const float t = lod - (float)l0;
  float4 acc = sample_level<DT>(A, S, p, ad, mx, my, mz, l0);
  if (t > 0.0f) acc = (1.0f - t) * acc + t * sample_level<DT>(A, S, p, ad, mx, my, mz, l0 + 1);
  return acc;
}

__device__ __forceinline__ float3 normalize3(float3 v) {
  const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return make_float3(v.x / l, v.y / l, v.z / l);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(float3 v) { return sqrtf(dot3(v, v)); }

// tangent (voxel_cone_tracing.frag:126-134)
__device__ __forceinline__ float3 tangent_of(float3 n) {
  const float3 t1 = cross3(n, make_float3(0.f, 0.f, 1.f));
  const float3 t2 = cross3(n, make_float3(0.f, 1.f, 0.f));
  return len3(t1) > len3(t2) ? normalize3(t1) : normalize3(t2);
}

__device__ __forceinline__ float3 refract3(float3 I, float3 N, float eta) {
  const float d = dot3(N, I);
  const float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return make_float3(0.f, 0.f, 0.f);
  const float s = eta * d + sqrtf(k);
  return make_float3(eta * I.x - s * N.x, eta * I.y - s * N.y, eta * I.z - s * N.z);
}

// One workgroup = a block of Shape::NPIX pixels (NSUB sub-blocks of 8x8) x 4
// waves.  Waves 0..NSUB-1 set the shading pixels of their sub-block up (lane =
// pixel) into LDS.  The cones of the block - every pixel's specular,
// refraction and shadow cones, then its diffuse cones, longest kinds first -
// form one queue of (job, pixel) items, job-major so that neighbouring lanes
// start on the same cone of neighbouring pixels (coherent gathers).  Every
// lane marches its own item and takes the next one from a workgroup counter
// in LDS as soon as its cone ends, so lanes do not idle while the longest cone
// of a wave finishes.  The same waves then compose their pixels in the
// reference's order (voxel_cone_tracing.frag:246-275).
// job codes: 0..15 diffuse cone k, then:
constexpr int kJobSpecular = 16, kJobRefraction = 17, kJobShadow = 18;  // kJobShadow + light
constexpr int kMaxJobs = kJobShadow + VCT_MAX_LIGHTS;
// idle lanes of a wave that trigger a refill: a refill costs the whole wave
// its set-up; measured on the 1M-triangle soup with 128-pixel blocks: 8 ->
// 2.69 ms, 16 -> 2.60 ms, 24 -> 2.55 ms, 32 -> 2.59 ms, 48 -> 2.77 ms
constexpr int kRefillMin = 24;

// Block shapes: pixels per workgroup and the shadow-light slots in LDS.
// Cones are short in cluttered scenes (6.2 steps per cone of a covered pixel
// on a 1M-triangle soup), so a queue of 64 pixels x 10 cones gives each of the 256 lanes only
// 2.5 items and the drain at its end leaves a third of the lanes idle; more
// pixels per workgroup shorten the drain relative to the work.  The wide
// shape holds shadow results of up to 2 lights; more lights use the full one.
template <int NPIX_, int NSHADOW_>
struct Shape {
  static constexpr int NPIX = NPIX_, NSHADOW = NSHADOW_, NSUB = NPIX_ / 64;
  static constexpr int BX = NSUB >= 2 ? 2 : 1;  // sub-blocks side by side (BX x NSUB/BX of 8x8)
};
using ShapeWide = Shape<128, 2>;
using ShapeFull = Shape<128, VCT_MAX_LIGHTS>;
using ShapeSmall = Shape<64, 2>;  // small viewports (see launch_trace)

// 32.32 fixed point of a non-negative cone colour channel (< 2^31): sums of
// these are exact, so the diffuse cones of a pixel may finish in any order.
__device__ __forceinline__ unsigned long long fix32(float v) { return __float2ull_rn(v * 4294967296.0f); }

// LDS of one workgroup of the cone marcher.
template <class SH>
struct ShadeLds {
  unsigned long long dsum[3][SH::NPIX];  // sum of the diffuse cones' rgb per pixel (fix32)
  float cone[2][3][SH::NPIX];            // specular, refraction rgb per pixel
  float shadow[SH::NSHADOW][SH::NPIX];   // max(0, 1 - alpha) of the shadow cone per light and pixel
  float pos[3][SH::NPIX], nrm[3][SH::NPIX], o1[3][SH::NPIX], o2[3][SH::NPIX], vdir[3][SH::NPIX];
  float spec_ap[SH::NPIX], eta[SH::NPIX];
  unsigned long long refr_mask[SH::NSUB];  // shading pixels with a transmissive material, per sub-block
  int sub_count[SH::NSUB];                 // shading pixels per sub-block
  // each set-up thread's G-buffer index (~0: none) and output index, kept
  // here over the cone queue instead of in registers
  uint32_t gidx[SH::NPIX], oidx[SH::NPIX];
  int npx, njobs, next_item;
  uint8_t px_of[SH::NPIX];  // the block's shading pixels, compacted
  int8_t jobs[kMaxJobs];
};

// The march state of one lane's cone (trace_cone, voxel_cone_tracing.frag:88-119).
// Registers are what bounds the marcher's occupancy, so only what cannot be
// derived is kept: |dir| and the three maps a sample reads follow from dir
// (abs is a free operand modifier), the diameter is dist * aperture, and the
// origin is the pixel's surface point, read from LDS where it is needed.
struct Cone {
  float3 dir;
  float ap, max_dist, dist;
  float4 c;
};

// The origin of the cone of block pixel `pix`, read from LDS where it is
// used.  A volatile read (forcing a re-read at every step) measured 0.11 ms
// slower on the 2560x1440 frame than letting the compiler schedule the reads.
template <int NPIX>
__device__ __forceinline__ float3 cone_origin(const float (*pos)[NPIX], int pix) {
  const float* p = &pos[0][0];
  return make_float3(p[pix], p[NPIX + pix], p[2 * NPIX + pix]);
}

__device__ __forceinline__ void cone_begin(Cone& k, float3 dir, float aperture, float max_dist, float vs) {
  dir = normalize3(dir);
  k.dir = dir;
  // A cone never needs to leave the volume by more than its diagonal; the
  // bound also keeps a non-finite target distance from looping forever.
  k.max_dist = fminf(max_dist, 3.4641016f);
  k.ap = aperture;
  k.c = f4(0.0f);
  k.dist = 3.0f * vs;
}

// The loop condition of trace_cone, plus the exact early exit: a sample more
// than half a texel of the coarsest level outside the unit cube reads only
// the zero border, so once the cone is that far out and moving away on that
// axis every later sample is zero and c no longer changes.
__device__ __forceinline__ bool cone_alive(const Cone& k, float3 origin, float margin) {
  if (!(k.c.w < 1.0f && k.dist < k.max_dist)) return false;
  const float3 sp = make_float3(k.dir.x * k.dist + origin.x, k.dir.y * k.dist + origin.y,
                                k.dir.z * k.dist + origin.z);
  return !((sp.x < -margin && k.dir.x <= 0.0f) || (sp.x >= 1.0f + margin && k.dir.x >= 0.0f) ||
           (sp.y < -margin && k.dir.y <= 0.0f) || (sp.y >= 1.0f + margin && k.dir.y >= 0.0f) ||
           (sp.z < -margin && k.dir.z <= 0.0f) || (sp.z >= 1.0f + margin && k.dir.z >= 0.0f));
}

template <int DT>
__device__ __forceinline__ void cone_step(const TraceArgs& A, const SamplerLds* S, Cone& k, float3 origin, float vs,
                                          float fres) {
  const float3 sp = make_float3(k.dir.x * k.dist + origin.x, k.dir.y * k.dist + origin.y,
                                k.dir.z * k.dist + origin.z);
  const float diam = k.dist * k.ap;
  const float mip = fmaxf(log2f(diam * fres), 0.0f);
  // face_indices (voxel_cone_tracing.frag:71-78): the map of each axis by the sign of dir
  const float3 ad = make_float3(fabsf(k.dir.x), fabsf(k.dir.y), fabsf(k.dir.z));
  const float4 s = sample_voxel<DT>(A, S, sp, ad, k.dir.x < 0.0f ? 0 : 1, k.dir.y < 0.0f ? 2 : 3,
                                    k.dir.z < 0.0f ? 4 : 5, mip);
  k.c = k.c + (1.0f - k.c.w) * s;
  k.dist += fmaxf(diam * 0.5f, vs);
}

// Direction of light li seen from pos (biased coordinates) and the distance
// the shadow cone may travel (direct_light, voxel_cone_tracing.frag:175-218).
__device__ __forceinline__ float3 light_dir(const TraceArgs& A, int li, float3 pos, float& d) {
  const DevLight& Lt = A.lights.l[li];
  if (Lt.kind == VCT_LIGHT_DIRECTIONAL) {
    d = 1.73205080757f;
    return normalize3(make_float3(Lt.pos[0], Lt.pos[1], Lt.pos[2]));
  }
  const float3 lp = make_float3(0.5f * (Lt.pos[0] / A.cube_size) + 0.5f, 0.5f * (Lt.pos[1] / A.cube_size) + 0.5f,
                                0.5f * (Lt.pos[2] / A.cube_size) + 0.5f);
  const float3 v = make_float3(lp.x - pos.x, lp.y - pos.y, lp.z - pos.z);
  d = len3(v);
  return normalize3(v);
}

template <int N>
__device__ __forceinline__ float3 ld3(const float (*a)[N], int p) { return make_float3(a[0][p], a[1][p], a[2][p]); }

// Shades the block's pixels; (px, py, oidx, valid) are the thread's pixel in
// sub-block (wave index mod NSUB).  L->next_item must be 0 and visible to the
// workgroup on entry.
template <int DT, class SH>
__device__ __forceinline__ void shade_block(const TraceArgs& A, int px, int py, size_t oidx, bool valid,
                                            ShadeLds<SH>* L, const SamplerLds* S) {
  const int lane = threadIdx.x & 63;
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform: a scalar register
  const int me = grp * 64 + lane;  // the thread's pixel in the block (set-up and composition waves)
  valid = valid && px < A.W && py < A.H;
  size_t idx = valid ? (size_t)py * A.W + px : 0;
  const float sqrt3 = 1.73205080757f;
  const int nd = A.en_diffuse ? A.cones.n : 0;
  const bool do_spec = A.en_specular != 0;
  const int nl = A.lights.count < VCT_MAX_LIGHTS ? A.lights.count : VCT_MAX_LIGHTS;
  // ---- set-up of the block's pixels (wave 0, lane = pixel) -------------------
  auto pixel = [&](float4& gp, float4& gn, const DevMaterial*& m, float3& pos, int& mode) {
    gp = A.gpos[idx];
    gn = A.gnrm[idx];
    uint32_t mid = __builtin_bit_cast(uint32_t, gn.w);
    mid = mid < A.nmat ? mid : 0u;  // ids of an uploaded G-buffer are not trusted
    m = &A.mats[mid];
    pos = make_float3(0.5f * (gp.x / A.cube_size) + 0.5f, 0.5f * (gp.y / A.cube_size) + 0.5f,
                      0.5f * (gp.z / A.cube_size) + 0.5f);
    // 0: nothing to do, 1: background / outside the cube, 2: voxel view, 3: shade
    mode = 0;
    if (valid) {
      if (gp.w == 0.0f || !(fabsf(pos.x) < 1.0f && fabsf(pos.y) < 1.0f && fabsf(pos.z) < 1.0f)) mode = 1;
      else mode = A.view_dir < 7 ? 2 : 3;  // within_cube on biased coordinates (:250, :61-64)
    }
  };
  unsigned long long shade = 0ull;
  int smode = 0;
  if (grp < SH::NSUB) {
    L->gidx[me] = valid ? (uint32_t)idx : ~0u;
    L->oidx[me] = (uint32_t)oidx;
    float4 gp, gn;
    const DevMaterial* m;
    float3 pos;
    int mode;
    pixel(gp, gn, m, pos, mode);
    smode = mode;
    const float3 n = make_float3(gn.x, gn.y, gn.z);
    const float3 cam = make_float3(A.cam[0], A.cam[1], A.cam[2]);
    const float3 vd = normalize3(make_float3(gp.x - cam.x, gp.y - cam.y, gp.z - cam.z));
    // specular aperture (:220-227)
    const float rough = sqrtf(2.0f / (m->shininess + 2.0f));
    L->spec_ap[me] = fminf(fmaxf(tanf(1.57079f * rough), 0.0174533f), 3.14159265f);
    L->eta[me] = 1.0f / m->ior;
    float3 o1 = make_float3(0.f, 0.f, 0.f), o2 = o1;
    if (mode == 3 && nd > 0) {
      o1 = normalize3(tangent_of(n));  // trace_diffuse (:146-147)
      o2 = normalize3(cross3(o1, n));
    }
    L->dsum[0][me] = 0ull;
    L->dsum[1][me] = 0ull;
    L->dsum[2][me] = 0ull;
    const float3 vals[5] = {pos, n, o1, o2, vd};
    float (*dst[5])[SH::NPIX] = {L->pos, L->nrm, L->o1, L->o2, L->vdir};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      dst[q][0][me] = vals[q].x;
      dst[q][1][me] = vals[q].y;
      dst[q][2][me] = vals[q].z;
    }
    shade = __ballot(mode == 3);
    const unsigned long long refr = __ballot(mode == 3 && do_spec && illum_transmissive(m->illum));
    if (lane == 0) {
      L->sub_count[grp] = __popcll(shade);
      L->refr_mask[grp] = refr;
    }
  }
  __syncthreads();
  if (grp < SH::NSUB) {
    // the block's shading pixels, compacted in sub-block order
    int base = 0, total = 0;
    bool refr = false;
#pragma unroll
    for (int w = 0; w < SH::NSUB; ++w) {
      base += w < grp ? L->sub_count[w] : 0;
      total += L->sub_count[w];
      refr = refr || L->refr_mask[w] != 0ull;
    }
    if (smode == 3) L->px_of[base + __popcll(shade & ((1ull << lane) - 1ull))] = (uint8_t)me;
    if (me == 0) {
      L->npx = total;
      int nj = 0;  // longest kinds first
      if (do_spec) L->jobs[nj++] = kJobSpecular;
      if (refr) L->jobs[nj++] = kJobRefraction;
      if (A.en_direct && A.en_shadow)
        for (int li = 0; li < nl; ++li) L->jobs[nj++] = (int8_t)(kJobShadow + li);
      for (int k = 0; k < nd; ++k) L->jobs[nj++] = (int8_t)k;
      L->njobs = nj;
    }
  }
  __syncthreads();
  // ---- the cone queue (every lane) -----------------------------------------------
  const float vs = 1.0f / (float)A.res, fres = (float)A.res;
  const float margin = 0.5f * (float)(1u << (A.levels - 1)) * vs;
  const int npx = L->npx;
  const int nitems = npx * L->njobs;
  // it / npx as (it * ceil(2^20 / npx)) >> 20: exact for it < kMaxJobs * 128 and npx <= 128
  const uint32_t npx_inv = npx ? (1u << 20) / (uint32_t)npx + (((1u << 20) % (uint32_t)npx) ? 1u : 0u) : 0u;
  // the lane's cone: job << 8 | pixel while it marches, -1 while it idles
  int cur = -1;
  unsigned long long wave_steps = 0;  // cone steps of the wave (uniform)
  bool drained = nitems == 0;
  Cone k;
  auto finish = [&](int job, int pix) {
    if (job >= kJobShadow) {
      L->shadow[job - kJobShadow][pix] = fmaxf(0.0f, 1.0f - k.c.w);
    } else {
      if (job < kJobSpecular) {
        atomicAdd(&L->dsum[0][pix], fix32(k.c.x));
        atomicAdd(&L->dsum[1][pix], fix32(k.c.y));
        atomicAdd(&L->dsum[2][pix], fix32(k.c.z));
      } else {
        const int slot = job == kJobSpecular ? 0 : 1;
        L->cone[slot][0][pix] = k.c.x;
        L->cone[slot][1][pix] = k.c.y;
        L->cone[slot][2][pix] = k.c.z;
      }
    }
  };
  for (;;) {
    // lanes without a cone take the next items of the queue, once enough of
    // them idle (or none is busy): a refill costs the whole wave its set-up
    const unsigned long long busy = __ballot(cur >= 0);
    for (;;) {
      const bool need = cur < 0 && !drained;
      const unsigned long long want = __ballot(need);
      if (!want || (busy && __popcll(want) < kRefillMin)) break;
      const int first = __ffsll((long long)want) - 1;
      int base = 0;
      if (lane == first) base = atomicAdd(&L->next_item, __popcll(want));
      base = __shfl(base, first);
      if (!need) continue;
      const int it = base + __popcll(want & ((1ull << lane) - 1ull));
      if (it >= nitems) {
        drained = true;
        continue;
      }
      const int j = (int)(((uint32_t)it * npx_inv) >> 20);
      const int job = L->jobs[j];
      const int pix = L->px_of[it - j * npx];
      if (job == kJobRefraction && !((L->refr_mask[pix >> 6] >> (pix & 63)) & 1ull)) continue;  // opaque: no item
      const float3 pos = ld3(L->pos, pix), n = ld3(L->nrm, pix);
      float3 d;
      float ap, md = sqrt3;
      if (job < kJobSpecular) {
        const float a = A.cones.a[job], b = A.cones.b[job], cc = A.cones.c[job];
        const float3 o1 = ld3(L->o1, pix), o2 = ld3(L->o2, pix);
        d = make_float3(a * n.x + b * o1.x + cc * o2.x, a * n.y + b * o1.y + cc * o2.y, a * n.z + b * o1.z + cc * o2.z);
        ap = A.cones.aperture;
      } else if (job == kJobSpecular) {
        // normalize(reflect(-view_dir, n)) (:231), normalised again in trace_cone (:90)
        const float3 vd = ld3(L->vdir, pix);
        const float3 I = make_float3(-vd.x, -vd.y, -vd.z);
        const float dn = 2.0f * dot3(n, I);
        d = normalize3(make_float3(I.x - dn * n.x, I.y - dn * n.y, I.z - dn * n.z));
        ap = L->spec_ap[pix];
      } else if (job == kJobRefraction) {
        d = refract3(ld3(L->vdir, pix), n, L->eta[pix]);  // trace_refraction (:237-241)
        ap = L->spec_ap[pix];
      } else {
        d = light_dir(A, job - kJobShadow, pos, md);  // shadow cone of direct_light
        ap = 0.1f;
      }
      cone_begin(k, d, ap, md, vs);
      if (cone_alive(k, pos, margin)) cur = (job << 8) | pix;
      else finish(job, pix);  // no step at all
    }
    const unsigned long long active = __ballot(cur >= 0);
    if (!active) break;
    wave_steps += (unsigned long long)__popcll(active);
    if (cur >= 0) {
      cone_step<DT>(A, S, k, cone_origin(L->pos, cur & 255), vs, fres);
      if (!cone_alive(k, cone_origin(L->pos, cur & 255), margin)) {
        finish(cur >> 8, cur & 255);
        cur = -1;
      }
    }
  }
  if (A.step_counter && lane == 0 && wave_steps) atomicAdd(A.step_counter, wave_steps);
  __syncthreads();
  // ---- composition (waves 0..NSUB-1, lane = pixel) -----------------------------------
  if (grp >= SH::NSUB) return;
  // the thread's slot again, from the scalar wave index and the hardware lane
  // id: nothing per lane is held over the cone queue
  const int lane_c = (int)__lane_id(), me_c = grp * 64 + lane_c;
  {
    const uint32_t g = ((const volatile uint32_t*)L->gidx)[me_c];
    valid = g != ~0u;
    idx = valid ? (size_t)g : 0;
    oidx = ((const volatile uint32_t*)L->oidx)[me_c];
  }
  float4 gp, gn;
  const DevMaterial* mp;
  float3 pos;
  int mode;
  pixel(gp, gn, mp, pos, mode);
  if (mode == 0) return;
  const float4 bg = make_float4(0.15f, 0.25f, 0.25f, 1.0f);  // renderer.cpp:398
  if (mode == 1) {
    A.image[oidx] = A.demodulate ? f4(0.0f) : bg;
    return;
  }
  if (mode == 2) {  // voxel view (:253-257), blended SRC_ALPHA / ONE_MINUS_SRC_ALPHA
    const int d = A.view_dir < 0 ? 0 : (A.view_dir > 5 ? 5 : A.view_dir);
    float3 ad = make_float3(0.f, 0.f, 0.f);
    int mx = 0, my = 2, mz = 4;
    if (d < 2) { ad.x = 1.0f; mx = d; }
    else if (d < 4) { ad.y = 1.0f; my = d; }
    else { ad.z = 1.0f; mz = d; }
    const float4 s = sample_voxel<DT>(A, S, pos, ad, mx, my, mz, A.view_lod);
    const float a = s.w;
    A.image[oidx] = make_float4(s.x * a + bg.x * (1.0f - a), s.y * a + bg.y * (1.0f - a),
                                s.z * a + bg.z * (1.0f - a), a * a + bg.w * (1.0f - a));
    return;
  }
  const DevMaterial& m = *mp;
  const float3 n = make_float3(gn.x, gn.y, gn.z);
  const float3 view_dir = ld3(L->vdir, me_c);
  const bool do_refr = (L->refr_mask[grp] >> lane_c) & 1ull;
  float3 out = make_float3(0.f, 0.f, 0.f);
  if (nd > 0) {
    // the exact sum of the diffuse cones, rounded once (trace_diffuse adds
    // them in cone order in FP32, :144-167)
    const float u = 1.0f / 4294967296.0f;
    const float3 diff_acc = make_float3(__ull2float_rn(L->dsum[0][me_c]) * u, __ull2float_rn(L->dsum[1][me_c]) * u,
                                        __ull2float_rn(L->dsum[2][me_c]) * u);
    const float inv = 1.0f / (float)nd;
    out.x += m.diffuse[0] * (diff_acc.x * inv);
    out.y += m.diffuse[1] * (diff_acc.y * inv);
    out.z += m.diffuse[2] * (diff_acc.z * inv);
  }
  if (A.en_direct) {  // direct_light (:175-218) with the shadow cones' results
    float3 res = make_float3(0.f, 0.f, 0.f);
    for (int li = 0; li < nl; ++li) {
      const DevLight& Lt = A.lights.l[li];
      float d;
      const float3 ldir = light_dir(A, li, pos, d);
      const float cos_surf = fmaxf(dot3(n, ldir), 0.0f);
      const float att = (Lt.kind == VCT_LIGHT_DIRECTIONAL) ? 1.0f : 1.0f / (1.0f + d * d);
      const float kk = att * cos_surf;
      const float3 lc = make_float3(kk * Lt.color[0] * Lt.intensity, kk * Lt.color[1] * Lt.intensity,
                                    kk * Lt.color[2] * Lt.intensity);
      const float shadow = A.en_shadow ? L->shadow[li][me_c] : 1.0f;
      const float lambert = fmaxf(dot3(ldir, n), 0.0f);
      float refr_angle = 0.0f;
      if (m.dissolve <= 0.1f) {
        const float3 r = refract3(view_dir, n, 1.0f / m.ior);
        refr_angle = fmaxf((1.0f - m.dissolve) * dot3(r, ldir), 0.0f);
      }
      const float3 hv = normalize3(make_float3(ldir.x + view_dir.x, ldir.y + view_dir.y, ldir.z + view_dir.z));
      float sa = fmaxf(dot3(hv, n), 0.0f);
      sa = fmaxf(sa, refr_angle);
      const float sc = powf(sa, m.shininess);
      const float f = shadow + 0.04f;
      res.x += f * (lambert * m.diffuse[0] + sc * m.specular[0]) * lc.x;
      res.y += f * (lambert * m.diffuse[1] + sc * m.specular[1]) * lc.y;
      res.z += f * (lambert * m.diffuse[2] + sc * m.specular[2]) * lc.z;
    }
    out.x += res.x + fminf(fmaxf(m.emission[0], 0.0f), 1.0f);
    out.y += res.y + fminf(fmaxf(m.emission[1], 0.0f), 1.0f);
    out.z += res.z + fminf(fmaxf(m.emission[2], 0.0f), 1.0f);
  }
  if (do_spec) {
    out.x += m.specular[0] * L->cone[0][0][me_c];
    out.y += m.specular[1] * L->cone[0][1][me_c];
    out.z += m.specular[2] * L->cone[0][2][me_c];
  }
  if (do_refr) {  // final colour override (:273-274)
    const float3 tr = make_float3(m.transmittance[0] * L->cone[1][0][me_c],
                                  m.transmittance[1] * L->cone[1][1][me_c],
                                  m.transmittance[2] * L->cone[1][2][me_c]);
    const float dv = m.dissolve;
    out = make_float3(tr.x * (1.0f - dv) + out.x * dv, tr.y * (1.0f - dv) + out.y * dv,
                      tr.z * (1.0f - dv) + out.z * dv);
  }
  float ow = 1.0f;
  if (A.demodulate) {
    const float e = 1.0f / 1024.0f;
    out = make_float3(out.x / fmaxf(m.diffuse[0], e), out.y / fmaxf(m.diffuse[1], e), out.z / fmaxf(m.diffuse[2], e));
    ow = gp.w;  // linear depth of the traced pixel
  }
  A.image[oidx] = make_float4(out.x, out.y, out.z, ow);
}

// Work of a launch: the whole viewport (trace_frame_kernel, one workgroup per
// block of NSUB 8x8 sub-blocks side by side), or a device-side list of pixels
// (NPIX per chunk) or tiles (16 sub-blocks of 8x8 pixels per tile, or 4 of
// 8x8 odd-odd grid points, NSUB per chunk) for trace_list_kernel.  5 waves per
// SIMD (96 VGPRs; the 8- and 16-bit storage paths fit without scratch, the
// FP32 path keeps 20 bytes per lane in scratch at this occupancy): the 12
// paired loads of a three-map level are in flight together.
template <int DT, class SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void trace_frame_kernel(TraceArgs A) {
  __shared__ ShadeLds<SH> s_shade;
  __shared__ SamplerLds s_sampler;
  const int lane = threadIdx.x & 63, sub = (int)(threadIdx.x >> 6) % SH::NSUB;
  sampler_init<DT>(A, &s_sampler);
  if (threadIdx.x == 0) s_shade.next_item = 0;
  __syncthreads();
  // Workgroups are dealt to the 8 XCDs round-robin in dispatch order.  XCD k
  // takes the block rows k, k + 8, k + 16, ... each left to right, so each
  // XCD's L2 serves whole rows of neighbouring blocks, and the 8 XCDs work
  // on rows spread over the image (rows left over below a multiple of 8
  // keep the dispatch order).  Measured at 2560x1440: trace 2.54 ms with
  // the dispatch order, 2.46 with one band of the image per XCD, 2.36 with
  // interleaved rows.
  const uint32_t gx = gridDim.x, gy = gridDim.y;
  uint32_t b = blockIdx.x + blockIdx.y * gx;
  if (b < (gy & ~7u) * gx) {
    const uint32_t q = b >> 3;
    b = ((q / gx) * 8u + (b & 7u)) * gx + q % gx;
  }
  const int bx = (int)(b % gx), by = (int)(b / gx);
  const int px = bx * (8 * SH::BX) + (sub % SH::BX) * 8 + (lane & 7);
  const int py = by * (8 * (SH::NSUB / SH::BX)) + (sub / SH::BX) * 8 + (lane >> 3);
  shade_block<DT, SH>(A, px, py, (size_t)py * A.W + px, true, &s_shade, &s_sampler);
}

// Chunks (npix pixels each) of a list launch.
__device__ __forceinline__ uint32_t list_chunks(const TraceArgs& A, uint32_t npix) {
  if (A.pix_list) return ((A.pix_count ? min(*A.pix_count, A.npix_list) : A.npix_list) + npix - 1u) / npix;
  return *A.tile_count * (A.grid_stride == 2 ? 4u : 16u) / (npix / 64u);
}

// The pixel of `lane` of sub-block `sub` in chunk c (nsub sub-blocks of 64
// pixels per chunk), where its result goes, and whether it is traced.
__device__ __forceinline__ void list_chunk(const TraceArgs& A, uint32_t c, uint32_t nsub, uint32_t sub, int lane,
                                           int& px, int& py, size_t& oidx, bool& valid) {
  valid = true;
  if (A.pix_list) {
    const uint32_t cnt = A.pix_count ? min(*A.pix_count, A.npix_list) : A.npix_list;
    const uint32_t k = (c * nsub + sub) * 64u + (uint32_t)lane;
    valid = k < cnt;
    const uint32_t p = valid ? A.pix_list[k] : 0u;
    px = (int)(p % (uint32_t)A.W);
    py = (int)(p / (uint32_t)A.W);
    oidx = A.pix_scatter ? (size_t)p : (size_t)k;
    if (valid && A.tile_flag) valid = A.tile_flag[(py >> 5) * A.tiles_x + (px >> 5)] == A.flag_want;
    return;
  }
  const uint32_t per = A.grid_stride == 2 ? 4u : 16u;  // sub-blocks per tile
  const uint32_t g = c * nsub + sub;                    // sub-block index over the list
  const uint32_t t = A.tile_list[g / per], b = g % per;
  if (A.grid_stride == 2) {  // grid point (u, v) is pixel (2u + 1, 2v + 1)
    const int gu = (int)(t % (uint32_t)A.tiles_x) * 16 + (int)(b & 1u) * 8 + (lane & 7);
    const int gv = (int)(t / (uint32_t)A.tiles_x) * 16 + (int)(b >> 1) * 8 + (lane >> 3);
    px = 2 * gu + 1;
    py = 2 * gv + 1;
    oidx = (size_t)gv * A.half_w + gu;
  } else {
    px = (int)(t % (uint32_t)A.tiles_x) * 32 + (int)(b & 3u) * 8 + (lane & 7);
    py = (int)(t / (uint32_t)A.tiles_x) * 32 + (int)(b >> 2) * 8 + (lane >> 3);
    oidx = (size_t)py * A.W + px;
  }
}

// One chunk per workgroup; workgroups past the device-side count leave at once.
// (A fixed grid striding over the chunks measured slower: the loop costs the
// kernel registers and spills at 4 waves per SIMD.)
template <int DT, class SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void trace_list_kernel(TraceArgs A) {
  __shared__ ShadeLds<SH> s_shade;
  __shared__ SamplerLds s_sampler;
  if (blockIdx.x >= list_chunks(A, SH::NPIX)) return;
  sampler_init<DT>(A, &s_sampler);
  if (threadIdx.x == 0) s_shade.next_item = 0;
  __syncthreads();
  int px, py;
  size_t oidx;
  bool valid;
  list_chunk(A, blockIdx.x, SH::NSUB, (threadIdx.x >> 6) % SH::NSUB, threadIdx.x & 63, px, py, oidx, valid);
  shade_block<DT, SH>(A, px, py, oidx, valid, &s_shade, &s_sampler);
}

// Builds the x-paired copy of every (level, map) of the volume in one launch:
// thread g covers texel g of the concatenated maps (level 0, then the 6 maps
// of each higher level).  The +x neighbour of Morton index m is the Morton
// increment of its x bits; the last column pairs with zero.
struct PairArgs {
  const char* src;
  char* dst;
  uint64_t src_off[16][6], dst_off[16][6];
  int first, levels;  // levels first .. levels - 1
  uint32_t res;
  uint64_t total;  // texels over all their maps
};

template <int DT>
__global__ __launch_bounds__(256) void pair_kernel(PairArgs P) {
  using T = typename Tex<DT>::Raw;
  using PR = typename Pair<DT>::Raw;
  for (uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x; g < P.total; g += (uint64_t)gridDim.x * 256u) {
    uint64_t r = g;
    int l = P.first, d = 0;
    uint32_t s = P.res >> l;
    uint64_t n = (uint64_t)s * s * s;
    if (l > 0 || r >= n) {
      if (l == 0) {
        r -= n;
        l = 1;
      }
      for (; l < P.levels; ++l) {
        s = P.res >> l;
        n = (uint64_t)s * s * s;
        if (r < 6 * n) break;
        r -= 6 * n;
      }
      d = (int)(r / n);
      r -= (uint64_t)d * n;
    }
    const uint32_t m = (uint32_t)r;
    const uint32_t mx = spread3(s - 1u);  // x bits of an index within this level
    const T* src = reinterpret_cast<const T*>(P.src + P.src_off[l][d]);
    const T a = src[m];
    T b;
    memset(&b, 0, sizeof(b));
    if ((m & mx) != mx) b = src[(((m | ~mx) + 1u) & mx) | (m & ~mx)];
    reinterpret_cast<PR*>(P.dst + P.dst_off[l][d])[m] = Pair<DT>::make(a, b);
  }
}

// Ceiling of the marcher's texel loads: the rate at which the vector memory
// path delivers 16-byte-per-lane loads (global_load_dwordx4, the form of the
// marcher's paired-texel gathers) when every access hits the L1 cache.  All
// workgroups read one 16 KiB region (L1 holds 32 KiB), each wave's lanes
// 1 KiB of consecutive addresses per load (8 full cache lines: the fewest
// lines a 64 x 16-byte load can touch), 8 loads of distinct rows in flight
// per lane; the row advances by a run-time stride, so no load can be merged
// with another or moved out of the loop.  A gather touches more lines per
// load, so no access pattern of these loads is delivered faster.
__global__ __launch_bounds__(256) void load_peak_kernel(const uint4* __restrict__ src, uint32_t iters,
                                                        uint32_t stride, uint4* __restrict__ sink) {
  uint4 acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = make_uint4(0u, 0u, 0u, 0u);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t i = 0; i < iters; i += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t row = (wave + i * stride + (uint32_t)u) & 15u;  // the 16 KiB region in 1 KiB rows
      const uint4 v = src[row * 64u + lane];
      acc[u].x ^= v.x;
      acc[u].y ^= v.y;
      acc[u].z ^= v.z;
      acc[u].w ^= v.w;
    }
  }
  uint4 r = acc[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) {
    r.x ^= acc[u].x;
    r.y ^= acc[u].y;
    r.z ^= acc[u].z;
    r.w ^= acc[u].w;
  }
  sink[(size_t)blockIdx.x * 256u + threadIdx.x] = r;  // keeps every load live
}

}  // namespace vct

using namespace vct;

extern "C" int vct_trace_load_peak(vct_ctx ctx, double* gbps) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  if (!gbps) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  VCT_HIP(hipSetDevice(ctx->device));
  int cus = 0;
  VCT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const uint32_t blocks = (uint32_t)std::max(cus, 1) * 8u, iters = 4096;
  int rc;
  VctBuf src, sink;
  if ((rc = vct_ensure(src, 16384, ctx))) return rc;
  if ((rc = vct_ensure(sink, (size_t)blocks * 256 * 16, ctx))) return rc;
  VCT_HIP(hipMemsetAsync(src.p, 0x5a, 16384, ctx->stream));
  float best = 0.0f;
  for (int rep = 0; rep < 4; ++rep) {  // the first launch also warms the caches
    VCT_HIP(hipEventRecord(ctx->ev[10], ctx->stream));
    hipLaunchKernelGGL(load_peak_kernel, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4*)src.p, iters, 1u,
                       (uint4*)sink.p);
    VCT_HIP(hipGetLastError());
    VCT_HIP(hipEventRecord(ctx->ev[11], ctx->stream));
    VCT_HIP(hipEventSynchronize(ctx->ev[11]));
    float ms = 0.0f;
    VCT_HIP(hipEventElapsedTime(&ms, ctx->ev[10], ctx->ev[11]));
    const double bytes = (double)blocks * 256.0 * iters * 16.0;
    const float g = (float)(bytes / (ms * 1e-3) / 1e9);
    if (rep > 0 && g > best) best = g;
  }
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(sink.p);
  (void)hipFree(src.p);
  *gbps = best;
  return VCT_OK;
}

// Cone direction tables in the tangent frame (n, o1, o2).
void vct_cone_set_table(int set, ConeSet& cs) {
  memset(&cs, 0, sizeof(cs));
  if (set == VCT_CONES_REF9) {
    // trace_diffuse (voxel_cone_tracing.frag:140-168): n, mix(n, +-o1/+-o2, .5),
    // mix(n, +-c1/+-c2, .5) with c1 = (o1 + o2) / 2, c2 = (o1 - o2) / 2.
    const float t[9][3] = {{1.f, 0.f, 0.f},      {.5f, .5f, 0.f},     {.5f, -.5f, 0.f},
                           {.5f, 0.f, .5f},      {.5f, 0.f, -.5f},    {.5f, .25f, .25f},
                           {.5f, -.25f, -.25f},  {.5f, .25f, -.25f},  {.5f, -.25f, .25f}};
    cs.n = 9;
    for (int i = 0; i < 9; ++i) {
      cs.a[i] = t[i][0];
      cs.b[i] = t[i][1];
      cs.c[i] = t[i][2];
    }
    cs.aperture = 0.55785173935f;  // TAN_22_5 (:136)
  } else if (set == VCT_CONES_5_WIDE) {
    const float t[5][3] = {{1.f, 0.f, 0.f}, {.5f, .5f, 0.f}, {.5f, -.5f, 0.f}, {.5f, 0.f, .5f}, {.5f, 0.f, -.5f}};
    cs.n = 5;
    for (int i = 0; i < 5; ++i) {
      cs.a[i] = t[i][0];
      cs.b[i] = t[i][1];
      cs.c[i] = t[i][2];
    }
    cs.aperture = 1.15470054f;  // 2 tan(30 deg): 60 degree cones
  } else {
    // 16 cones: the normal, 5 at 28 degrees, 10 at 58 degrees from it.
    cs.n = 16;
    cs.a[0] = 1.f;
    int k = 1;
    const double d2r = 3.14159265358979323846 / 180.0;
    for (int i = 0; i < 5; ++i, ++k) {
      double th = 28.0 * d2r, ph = 72.0 * i * d2r;
      cs.a[k] = (float)cos(th);
      cs.b[k] = (float)(sin(th) * cos(ph));
      cs.c[k] = (float)(sin(th) * sin(ph));
    }
    for (int i = 0; i < 10; ++i, ++k) {
      double th = 58.0 * d2r, ph = (36.0 * i + 18.0) * d2r;
      cs.a[k] = (float)cos(th);
      cs.b[k] = (float)(sin(th) * cos(ph));
      cs.c[k] = (float)(sin(th) * sin(ph));
    }
    cs.aperture = 0.45f;
  }
}

// Rebuilds the x-paired copy of the volume when the volume has changed since
// it was last built (16-bit and RGBA8 storage).  The mip pass calls it last;
// the marcher calls it again for volumes written since (uploads).
int vct_pair_layout(vct_context* ctx) {
  const uint64_t bpe = dtype_bytes(ctx->dtype);
  uint64_t bytes = 0;
  for (int l = 0; l < ctx->levels; ++l) {
    const uint64_t map_bytes = (level_elems(ctx->res >> l) * 2 * bpe + 255) & ~255ull;
    for (int d = 0; d < 6; ++d) {
      ctx->pair_off[l][d] = bytes;
      if (l > 0) bytes += map_bytes;
    }
    if (l == 0) bytes += map_bytes;
  }
  return vct_ensure(ctx->pairs, bytes, ctx);
}

int vct_pairs_from(vct_context* ctx, int first) {
  int rc;
  if ((rc = vct_pair_layout(ctx))) return rc;
  PairArgs P;
  memset(&P, 0, sizeof(P));
  uint64_t total = 0;
  for (int l = 0; l < ctx->levels; ++l) {
    for (int d = 0; d < 6; ++d) {
      P.src_off[l][d] = ctx->lvl_off[l][d];
      P.dst_off[l][d] = ctx->pair_off[l][d];
    }
    if (l >= first) total += (l == 0 ? 1 : 6) * level_elems(ctx->res >> l);
  }
  if (total == 0) return VCT_OK;
  P.src = (const char*)ctx->volume.p;
  P.dst = (char*)ctx->pairs.p;
  P.first = first;
  P.levels = ctx->levels;
  P.res = ctx->res;
  P.total = total;
  const dim3 grid((unsigned)std::min<uint64_t>((total + 255) / 256, 256ull * 64));
  switch (ctx->dtype) {
    case VCT_RGBA8: hipLaunchKernelGGL(pair_kernel<VCT_RGBA8>, grid, dim3(256), 0, ctx->stream, P); break;
    case VCT_FP16: hipLaunchKernelGGL(pair_kernel<VCT_FP16>, grid, dim3(256), 0, ctx->stream, P); break;
    default: hipLaunchKernelGGL(pair_kernel<VCT_BF16>, grid, dim3(256), 0, ctx->stream, P); break;
  }
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

int vct_ensure_pairs(vct_context* ctx) {
  if (ctx->dtype == VCT_FP32 || ctx->pairs_version == ctx->volume_version) return VCT_OK;
  int rc;
  if ((rc = vct_pairs_from(ctx, 0))) return rc;
  ctx->pairs_version = ctx->volume_version;
  return VCT_OK;
}

int vct_trace_into(vct_context* ctx, const float4* gpos, const float4* gnrm, float4* out, int W, int H,
                   int demodulate, unsigned long long* step_counter) {
  TraceSelect sel;
  memset(&sel, 0, sizeof(sel));
  return vct_trace_select(ctx, gpos, gnrm, out, W, H, demodulate, step_counter, sel);
}

// Launches the marcher of one block shape: the viewport, or one workgroup
// per chunk of the largest the list can be.
template <int DT, class SH>
static void launch_shape(vct_context* ctx, const TraceArgs& A, const TraceSelect& sel, int W, int H) {
  constexpr int bw = 8 * SH::BX, bh = 8 * (SH::NSUB / SH::BX);
  dim3 grid((W + bw - 1) / bw, (H + bh - 1) / bh);
  if (sel.pix_list) {
    grid = dim3((sel.npix_list + SH::NPIX - 1) / SH::NPIX);
    hipLaunchKernelGGL((trace_list_kernel<DT, SH>), grid, dim3(256), 0, ctx->stream, A);
  } else if (sel.tile_list) {
    grid = dim3((sel.grid_stride == 2 ? 4u : 16u) / SH::NSUB * sel.max_tiles);
    hipLaunchKernelGGL((trace_list_kernel<DT, SH>), grid, dim3(256), 0, ctx->stream, A);
  } else {
    hipLaunchKernelGGL((trace_frame_kernel<DT, SH>), grid, dim3(256), 0, ctx->stream, A);
  }
}

template <int DT>
static int launch_trace(vct_context* ctx, const TraceArgs& A, const TraceSelect& sel, int W, int H, bool wide) {
  if ((sel.pix_list && sel.npix_list == 0) || (!sel.pix_list && sel.tile_list && sel.max_tiles == 0)) return VCT_OK;
  // A viewport (or pixel list) of at most 2^18 pixels is at most 2048
  // workgroups of 128 pixels, under two rounds of the ~1280 that 256 CUs hold
  // at once, so the last round runs part-empty; 64-pixel workgroups halve
  // that tail.  Measured on the 1M-triangle soup: 640x360 0.384 -> 0.345 ms
  // (no change at 1280x720 and above, where 128 pixels shorten each
  // workgroup's drain instead); the learned frame's 57,600 probes beside it,
  // 3.257 -> 3.244 ms for the frame.
  if (wide && !sel.tile_list && (sel.pix_list ? (size_t)sel.npix_list : (size_t)W * H) <= (1u << 18))
    launch_shape<DT, ShapeSmall>(ctx, A, sel, W, H);
  else if (wide) launch_shape<DT, ShapeWide>(ctx, A, sel, W, H);
  else launch_shape<DT, ShapeFull>(ctx, A, sel, W, H);
  return VCT_OK;
}

int vct_trace_select(vct_context* ctx, const float4* gpos, const float4* gnrm, float4* out, int W, int H,
                     int demodulate, unsigned long long* step_counter, const TraceSelect& sel) {
  TraceArgs A;
  memset(&A, 0, sizeof(A));
  A.step_counter = step_counter;
  A.pix_list = sel.pix_list;
  A.npix_list = sel.npix_list;
  A.pix_count = sel.pix_count;
  A.pix_scatter = sel.pix_scatter;
  A.tile_list = sel.tile_list;
  A.tile_count = sel.tile_count;
  A.tiles_x = sel.tiles_x;
  A.tile_flag = sel.tile_flag;
  A.flag_want = sel.flag_want;
  A.grid_stride = sel.grid_stride;
  A.half_w = sel.half_w;
  int rc;
  if (ctx->res > (uint32_t)kMaxEdge) return vct_set_error(VCT_ERR_LIMIT, "the marcher samples grids of at most 512^3");
  if ((rc = vct_ensure_pairs(ctx))) return rc;
  const bool paired = ctx->dtype != VCT_FP32;
  A.vol = (const char*)(paired ? ctx->pairs.p : ctx->volume.p);
  for (int l = 0; l < ctx->levels; ++l)
    for (int d = 0; d < 6; ++d) A.off[l][d] = paired ? ctx->pair_off[l][d] : ctx->lvl_off[l][d];
  A.levels = ctx->levels;
  A.res = (int)ctx->res;
  A.gpos = gpos;
  A.gnrm = gnrm;
  A.mats = (const DevMaterial*)ctx->d_materials.p;
  A.nmat = (uint32_t)ctx->materials.size();
  A.image = out;
  A.W = W;
  A.H = H;
  A.cube_size = ctx->cube_size;
  // camera_position = column 3 of the view matrix (renderer.cpp:279)
  A.cam[0] = ctx->view[12];
  A.cam[1] = ctx->view[13];
  A.cam[2] = ctx->view[14];
  A.en_direct = ctx->en_direct;
  A.en_diffuse = ctx->en_diffuse;
  A.en_specular = ctx->en_specular;
  A.en_shadow = ctx->en_shadow;
  A.view_dir = ctx->view_voxel_dir;
  A.view_lod = ctx->view_voxel_lod;
  A.demodulate = demodulate;
  vct_cone_set_table(ctx->cone_set, A.cones);
  int n = (int)ctx->lights.size();
  if (n > VCT_MAX_LIGHTS) n = VCT_MAX_LIGHTS;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      A.lights.l[i].pos[k] = ctx->lights[i].position[k];
      A.lights.l[i].color[k] = ctx->lights[i].color[k];
    }
    A.lights.l[i].intensity = ctx->lights[i].intensity;
    A.lights.l[i].kind = ctx->lights[i].kind;
  }
  A.lights.count = n;
  // the wide shape holds the shadows of up to 2 lights
  const int nshadow = (ctx->en_direct && ctx->en_shadow) ? n : 0;
  const bool wide = nshadow <= ShapeWide::NSHADOW;
  switch (ctx->dtype) {
    case VCT_RGBA8: rc = launch_trace<VCT_RGBA8>(ctx, A, sel, W, H, wide); break;
    case VCT_FP16: rc = launch_trace<VCT_FP16>(ctx, A, sel, W, H, wide); break;
    case VCT_BF16: rc = launch_trace<VCT_BF16>(ctx, A, sel, W, H, wide); break;
    default: rc = launch_trace<VCT_FP32>(ctx, A, sel, W, H, wide); break;
  }
  if (rc) return rc;
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

int vct_launch_trace(vct_context* ctx) {
  const size_t npix = (size_t)ctx->width * ctx->height;
  if (!ctx->g_pos.p || ctx->g_pos.cap < npix * 16)
    return vct_set_error(VCT_ERR_STATE, "trace needs a G-buffer of the current viewport (run vct_gbuffer first)");
  int rc;
  if ((rc = vct_ensure(ctx->image, npix * 16, ctx))) return rc;
  // the voxel view blends every fragment in draw order (renderer.cpp:384-388);
  // from a rasterised G-buffer its triangles are known, from an uploaded one
  // only the surface is, which is then blended over the background
  if (ctx->view_voxel_dir < 7 && ctx->g_frame.valid && ctx->g_frame.W == ctx->width &&
      ctx->g_frame.H == ctx->height)
    return vct_launch_voxel_view(ctx);
  return vct_trace_into(ctx, (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p, (float4*)ctx->image.p,
                        ctx->width, ctx->height, 0, nullptr);
}

extern "C" int vct_trace_count_steps(vct_ctx ctx, uint64_t* steps) {
  if (!ctx) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");
  if (!steps) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  VCT_HIP(hipSetDevice(ctx->device));
  const size_t npix = (size_t)ctx->width * ctx->height;
  if (!ctx->g_pos.p || ctx->g_pos.cap < npix * 16)
    return vct_set_error(VCT_ERR_STATE, "trace needs a G-buffer of the current viewport (run vct_gbuffer first)");
  int rc;
  if ((rc = vct_ensure(ctx->image, npix * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->counters, 256, ctx))) return rc;
  unsigned long long* cnt = (unsigned long long*)((char*)ctx->counters.p + 64);
  VCT_HIP(hipMemsetAsync(cnt, 0, 8, ctx->stream));
  if ((rc = vct_trace_into(ctx, (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p, (float4*)ctx->image.p,
                           ctx->width, ctx->height, 0, cnt)))
    return rc;
  unsigned long long h = 0;
  VCT_HIP(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  *steps = h;
  return VCT_OK;
}
