// vct_mlp.hip - learned quarter-resolution upsampler (no reference counterpart).
//
// A per-pixel MLP 34 -> 64 -> 64 -> 3 with ReLU.  Master parameters are FP32
// (flat: W1[64][34], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]); every
// matrix product runs on the gfx950 matrix cores as
// v_mfma_f32_16x16x32_bf16 with BF16 operands and FP32 accumulation.
//
// Matrix-core tile helper: C[16x16] += A[16xK] * B[Kx16] where both operands
// live in LDS (or global memory) with K contiguous: A as rows m, B as rows n.
// Lane l supplies A[m = l%16][k0 + 8*(l/16) .. +7] and B[n = l%16][same k] as
// one 16-byte load each and receives C[4*(l/16) + i][l%16], i = 0..3.
//
// Features of a full-resolution pixel p (index order kappa):
//   0..26  the 3x3 neighbourhood of the quarter-resolution indirect light
//          (albedo-demodulated), bilinearly sampled at p's own position plus
//          offsets of -1, 0, +1 low-resolution pixels (tap-major, RGB)
//   27..29 normal, 30 relative linear depth (d - d_low) / d_low, 31..33 albedo
// The network predicts the demodulated indirect light; the image is
// albedo * prediction.
#include <cmath>
#include <cstring>
#include <mutex>

#include "vct_internal.h"

namespace vct {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kIn = VCT_MLP_IN, kHid = VCT_MLP_HIDDEN, kOut = VCT_MLP_OUT;
constexpr int kKP = 64;  // padded input width
constexpr int kFbTile = 32;  // the fallback's tiles (pixels per edge)
// flat master layout
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kOut * kHid, kParams = kOffB3 + kOut;
static_assert(kParams == VCT_MLP_PARAMS, "parameter count");

// Packed BF16 operand copies (element offsets in one buffer, all K contiguous)
//   W1p[64][64]  (out x in, in >= 34 zero)      A of layer 1
//   W2p[64][64]  (out x in)                     A of layer 2
//   W3p[16][64]  (out x in, out >= 3 zero)      A of layer 3
//   W2T[64][64]  (in x out)                     A of dH1 = W2^T dZ2
//   W3T[64][32]  (in x out, out >= 3 zero)      A of dH2 = W3^T dY
constexpr int kPW1 = 0, kPW2 = kPW1 + 64 * 64, kPW3 = kPW2 + 64 * 64, kPW2T = kPW3 + 16 * 64,
              kPW3T = kPW2T + 64 * 64, kPacked = kPW3T + 64 * 32;

__device__ __forceinline__ uint16_t bf16_bits(float f) { return f32_to_bf16(f); }

// One 16x16x32 MFMA step; a and b point at the lane's 8 contiguous BF16.
__device__ __forceinline__ floatx4 mfma_step(const uint16_t* a, const uint16_t* b, floatx4 c) {
  const bf16x8 va = *reinterpret_cast<const bf16x8*>(a);
  const bf16x8 vb = *reinterpret_cast<const bf16x8*>(b);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, c, 0, 0, 0);
}

// C tile += A[m0:m0+16][0:K] * B[n0:n0+16][0:K]^T, K a multiple of 32.
__device__ __forceinline__ floatx4 mfma_tile(const uint16_t* A, int lda, const uint16_t* B, int ldb, int K,
                                             floatx4 c) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += 32) c = mfma_step(A + r * lda + k0 + kq, B + r * ldb + k0 + kq, c);
  return c;
}

struct MlpDev {
  const uint16_t* packed;  // kPacked BF16
  const float* b1;
  const float* b2;
  const float* b3;
};

// ---- features ---------------------------------------------------------------------
struct FeatArgs {
  const float4* gpos;  // full-res G-buffer
  const float4* gnrm;
  const DevMaterial* mats;
  uint32_t nmat;
  const float4* low;   // quarter-res demodulated indirect light (rgb) + linear depth (w)
  const float4* lnrm;  // quarter-res normals + material id (the G-buffer at the traced pixels)
  int W, H, wl, hl;
};

enum : int { kBackground = 0, kFeatures = 1, kTraceIt = 2 };

// Features of full-resolution pixel (px, py).  Low-resolution pixel (i, j)
// was traced at full-resolution pixel (4i + 1, 4j + 1).  The 3x3 taps around
// the nearest low-resolution pixel are weighted by a tent of radius 2
// low-resolution pixels and by whether the tap lies on the same surface as
// the pixel (same material, normals within ~18 degrees, linear depths within
// 5 %); f[0..26] are the taps' RGB times their normalised weights, `base`
// their sum (the surface-guided interpolation the network corrects), f[27..29]
// the unit normal, f[30] the relative depth to the taps' weighted depth, and
// f[31..33] the albedo.  Returns kBackground, kFeatures, or kTraceIt when no
// tap lies on the pixel's surface (the pixel is traced instead).
__device__ int pixel_features(const FeatArgs& F, int px, int py, float* f, float* alb, float3& base) {
  const size_t idx = (size_t)py * F.W + px;
  const float4 gp = F.gpos[idx];
  if (gp.w == 0.0f) return kBackground;
  const float4 gn = F.gnrm[idx];
  const uint32_t mbits = __builtin_bit_cast(uint32_t, gn.w);
  const uint32_t mid = mbits < F.nmat ? mbits : 0u;
  const DevMaterial& m = F.mats[mid];
  const float nl = sqrtf(gn.x * gn.x + gn.y * gn.y + gn.z * gn.z);
  const float inl = nl > 0.0f ? 1.0f / nl : 0.0f;
  const float nx = gn.x * inl, ny = gn.y * inl, nz = gn.z * inl;
  alb[0] = m.diffuse[0];
  alb[1] = m.diffuse[1];
  alb[2] = m.diffuse[2];
  const float u = (float)(px - 1) * 0.25f, v = (float)(py - 1) * 0.25f;
  const int ic = (int)floorf(u + 0.5f), jc = (int)floorf(v + 0.5f);
  float w[9], wsum = 0.0f, dsum = 0.0f;
  float4 c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int i = ic + (k % 3) - 1, j = jc + (k / 3) - 1;
    w[k] = 0.0f;
    c[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < 0 || i >= F.wl || j < 0 || j >= F.hl) continue;
    const size_t q = (size_t)j * F.wl + i;
    const float4 ln = F.lnrm[q];
    const float4 lv = F.low[q];
    if (__builtin_bit_cast(uint32_t, ln.w) != mbits || !(lv.w > 0.0f)) continue;
    const float d = nx * ln.x + ny * ln.y + nz * ln.z;
    const float ll = sqrtf(ln.x * ln.x + ln.y * ln.y + ln.z * ln.z);
    if (!(d > 0.95f * ll) || !(fabsf(lv.w - gp.w) < 0.05f * gp.w)) continue;
    const float wk = fmaxf(0.0f, 1.0f - 0.5f * fabsf(u - (float)i)) * fmaxf(0.0f, 1.0f - 0.5f * fabsf(v - (float)j));
    w[k] = wk;
    c[k] = lv;
    wsum += wk;
    dsum += wk * lv.w;
  }
  if (!(wsum > 0.0f)) return kTraceIt;
  const float iw = 1.0f / wsum;
  base = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const float t = w[k] * iw;
    f[3 * k + 0] = c[k].x * t;
    f[3 * k + 1] = c[k].y * t;
    f[3 * k + 2] = c[k].z * t;
    base.x += f[3 * k + 0];
    base.y += f[3 * k + 1];
    base.z += f[3 * k + 2];
  }
  f[27] = nx;
  f[28] = ny;
  f[29] = nz;
  const float dw = dsum * iw;
  f[30] = fminf(fmaxf((gp.w - dw) / gp.w, -1.0f), 1.0f);
  f[31] = m.diffuse[0];
  f[32] = m.diffuse[1];
  f[33] = m.diffuse[2];
  return kFeatures;
}

// Row stride (BF16 elements) of the activations a wave keeps in LDS: 64
// values padded to 72, so the 16 rows one MFMA operand load reads start 16
// banks apart instead of all on the same bank (a 128-byte stride put the
// inference kernel at 30 LDS bank conflicts per LDS instruction).
constexpr int kXS = 72;

// ---- forward of one wave's 64 samples held in LDS ---------------------------------
// xs: [64 samples][kXS] BF16 (features zero padded to 64).  Layers 1 and 2 are
// computed per 16-sample column tile and written back over that tile's own
// input rows, which no other tile reads (the tile's MFMAs have consumed them
// before the write; one wave, in order).  Returns y (C layout of the 16-row
// output tile) for the wave's four 16-sample column tiles.
__device__ __forceinline__ void forward_wave(const MlpDev& P, uint16_t* xs, floatx4 (&y)[4]) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, rq = (lane >> 4) * 4;
  for (int layer = 0; layer < 2; ++layer) {
    const uint16_t* Wp = P.packed + (layer == 0 ? kPW1 : kPW2);
    const float* bias = layer == 0 ? P.b1 : P.b2;
    for (int nt = 0; nt < 4; ++nt) {
      floatx4 acc[4];
      for (int mt = 0; mt < 4; ++mt) {
        acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc[mt] = mfma_tile(Wp + mt * 16 * 64, 64, xs + nt * 16 * kXS, kXS, 64, acc[mt]);
      }
      for (int mt = 0; mt < 4; ++mt) {
        uint16_t o[4];
        for (int i = 0; i < 4; ++i) {
          const float h = fmaxf(acc[mt][i] + bias[mt * 16 + rq + i], 0.0f);
          o[i] = bf16_bits(h);
        }
        uint2 pk = make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
        *reinterpret_cast<uint2*>(xs + (nt * 16 + col) * kXS + mt * 16 + rq) = pk;
      }
    }
  }
  for (int nt = 0; nt < 4; ++nt) {
    floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(P.packed + kPW3, 64, xs + nt * 16 * kXS, kXS, 64, acc);
    for (int i = 0; i < 4; ++i) acc[i] += (rq + i < kOut) ? P.b3[rq + i] : 0.0f;
    y[nt] = acc;
  }
}

// ---- inference: features + MLP + output for a 16x16 pixel tile per workgroup ------
// Output = albedo * (guided interpolation + network correction).  Pixels with
// no same-surface tap are appended to `trace_list` (traced afterwards at full
// resolution); pixels that were traced at quarter resolution keep that value.
struct InferArgs {
  FeatArgs F;
  MlpDev P;
  float4* image;
  uint32_t* trace_list;  // [0]: count (zero before the launch), then pixel indices; null: not listed
  // optional: the pixels of a list (256 per workgroup) instead of 16x16 blocks
  const uint32_t* pix_list;
  uint32_t npix_list;
  // optional, blocks only: skip the 16x16 blocks of 32x32 tiles with skip_tile[t] == 1
  const uint32_t* skip_tile;
  int tiles_x;
};

// Each lane keeps its pixel's data in registers over the network and writes
// its own pixel, taking the three outputs of its sample from the lane that
// holds them (36 KB of LDS per workgroup: 4 workgroups per CU).
__global__ __launch_bounds__(256) void mlp_infer_kernel(InferArgs A) {
  __shared__ __attribute__((aligned(16))) uint16_t s_x[256 * kXS];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  if (!A.pix_list && A.skip_tile &&
      A.skip_tile[(blockIdx.y * 16 / kFbTile) * A.tiles_x + blockIdx.x * 16 / kFbTile] == 1u)
    return;  // the whole workgroup: its tile is traced in full
  int px = blockIdx.x * 16 + (t & 15), py = blockIdx.y * 16 + (t >> 4);
  if (A.pix_list) {
    const uint32_t k = blockIdx.x * 256u + (uint32_t)t;
    px = py = -1;
    if (k < A.npix_list) {
      const uint32_t q = A.pix_list[k];
      px = (int)(q % (uint32_t)A.F.W);
      py = (int)(q / (uint32_t)A.F.W);
    }
  }
  float f[kIn], alb[3] = {0.f, 0.f, 0.f};
  float3 base = make_float3(0.f, 0.f, 0.f);
  int st = kBackground;
  const bool inside = px >= 0 && py >= 0 && px < A.F.W && py < A.F.H;
  if (inside) st = pixel_features(A.F, px, py, f, alb, base);
  if (A.trace_list) {  // pixels to trace: one atomic per wave
    const unsigned long long m = __ballot(st == kTraceIt);
    uint32_t b0 = 0;
    if (lane == 0 && m) b0 = atomicAdd(A.trace_list, (uint32_t)__popcll(m));
    b0 = __shfl(b0, 0, 64);
    if (st == kTraceIt)
      A.trace_list[1 + b0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)(py * A.F.W + px);
  }
  uint16_t* row = s_x + t * kXS;
  const bool feat = st == kFeatures;
  for (int k = 0; k < 64; k += 8) {
    uint32_t w[4];
    for (int q = 0; q < 4; ++q) {
      const int k0 = k + 2 * q, k1 = k0 + 1;
      const uint32_t lo = (feat && k0 < kIn) ? bf16_bits(f[k0]) : 0u;
      const uint32_t hi = (feat && k1 < kIn) ? bf16_bits(f[k1]) : 0u;
      w[q] = lo | (hi << 16);
    }
    *reinterpret_cast<uint4*>(row + k) = make_uint4(w[0], w[1], w[2], w[3]);
  }
  __syncthreads();
  floatx4 y[4];
  forward_wave(A.P, s_x + wave * 64 * kXS, y);
  // sample t is column (lane & 15) of the wave's column tile lane >> 4; lanes
  // 0..15 hold rows 0..3 (outputs 0..2) of each tile's columns
  const int src = lane & 15, tile = lane >> 4;
  float out[3];
  for (int i = 0; i < 3; ++i) {
    const float v0 = __shfl(y[0][i], src, 64), v1 = __shfl(y[1][i], src, 64);
    const float v2 = __shfl(y[2][i], src, 64), v3 = __shfl(y[3][i], src, 64);
    out[i] = tile == 0 ? v0 : (tile == 1 ? v1 : (tile == 2 ? v2 : v3));
  }
  if (!inside) return;
  float4 o;
  if (st == kTraceIt) {  // left for tracing
    o = make_float4(0.f, 0.f, 0.f, -1.0f);
  } else if (st == kFeatures) {
    float3 v = make_float3(base.x + out[0], base.y + out[1], base.z + out[2]);
    if ((px & 3) == 1 && (py & 3) == 1 && (px >> 2) < A.F.wl && (py >> 2) < A.F.hl) {
      // this pixel was traced at quarter resolution: keep the traced value
      const float4 tv = A.F.low[(size_t)(py >> 2) * A.F.wl + (px >> 2)];
      v = make_float3(tv.x, tv.y, tv.z);
    }
    o = make_float4(alb[0] * v.x, alb[1] * v.y, alb[2] * v.z, 1.0f);
  } else {
    o = make_float4(0.15f, 0.25f, 0.25f, 1.0f);  // clear colour, renderer.cpp:398
  }
  A.image[(size_t)py * A.F.W + px] = o;
}

// ---- training samples: features, albedo and residual target at listed pixels ------
// Y = target - guided interpolation (the network learns the correction);
// background pixels and pixels without a same-surface tap get zero weight.
__global__ __launch_bounds__(256) void mlp_sample_kernel(FeatArgs F, const float4* __restrict__ target,
                                                         const uint32_t* __restrict__ pixels, uint32_t n,
                                                         uint16_t* __restrict__ X, float* __restrict__ Alb,
                                                         float* __restrict__ Y, uint32_t* __restrict__ valid) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t p = pixels[k];
  const int px = (int)(p % (uint32_t)F.W), py = (int)(p / (uint32_t)F.W);
  float f[kIn], alb[3] = {0.f, 0.f, 0.f};
  float3 base = make_float3(0.f, 0.f, 0.f);
  const bool ok = pixel_features(F, px, py, f, alb, base) == kFeatures;
  uint16_t* row = X + (size_t)k * 64;
  for (int c = 0; c < 64; ++c) row[c] = (ok && c < kIn) ? bf16_bits(f[c]) : (uint16_t)0;
  const float4 tg = target[p];
  Alb[k * 3 + 0] = ok ? alb[0] : 0.f;
  Alb[k * 3 + 1] = ok ? alb[1] : 0.f;
  Alb[k * 3 + 2] = ok ? alb[2] : 0.f;
  Y[k * 3 + 0] = ok ? tg.x - base.x : 0.f;
  Y[k * 3 + 1] = ok ? tg.y - base.y : 0.f;
  Y[k * 3 + 2] = ok ? tg.z - base.z : 0.f;
  valid[k] = ok ? 1u : 0u;
}

// ---- forward on given features (tests and evaluation) -------------------------------
__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpDev P, const uint16_t* __restrict__ X, uint32_t n,
                                                          float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) uint16_t s_x[256 * kXS];
  const int t = threadIdx.x;
  const uint32_t base = blockIdx.x * 256u;
  {
    const uint32_t k = base + t;
    const uint4* src = reinterpret_cast<const uint4*>(X + (size_t)k * 64);
    uint4* dst = reinterpret_cast<uint4*>(s_x + t * kXS);
    for (int q = 0; q < 8; ++q) dst[q] = k < n ? src[q] : make_uint4(0, 0, 0, 0);
  }
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  floatx4 y[4];
  forward_wave(P, s_x + wave * 64 * kXS, y);
  if (lane < 16)
    for (int nt = 0; nt < 4; ++nt) {
      const uint32_t k = base + wave * 64 + nt * 16 + lane;
      if (k < n)
        for (int o = 0; o < kOut; ++o) Y[(size_t)k * 3 + o] = y[nt][o];
    }
}

// ---- training: forward + backward over chunks of 64 samples ------------------------
// Per chunk (P = 64 samples, 4 waves of 16): every activation and gradient is
// staged in LDS as BF16 in the K-contiguous layout its next product reads.
// Gradients accumulate in FP32 matrix-core accumulators across the chunks of
// the workgroup; partial sums go to `partial` [gridDim.x][kParams] and are
// reduced in a fixed order by mlp_adam_kernel (deterministic).
struct TrainArgs {
  MlpDev P;
  const uint16_t* X;  // [n][64]
  const float* Alb;   // [n][3]
  const float* Y;     // [n][3] target
  uint32_t n;
  float scale;        // 2 / (3 * n): d(mean squared error) / dy
  float* partial;     // [gridDim.x][kParams + 1] (last: loss sum)
};

constexpr int kP = 64;

__global__ __launch_bounds__(256) void mlp_train_kernel(TrainArgs A) {
  __shared__ __attribute__((aligned(16))) uint16_t xs[kP * 64];    // [sample][in]
  __shared__ __attribute__((aligned(16))) uint16_t xt[64 * kP];    // [in][sample]
  __shared__ __attribute__((aligned(16))) uint16_t h1s[kP * 64];   // [sample][hid]
  __shared__ __attribute__((aligned(16))) uint16_t h1t[64 * kP];   // [hid][sample]
  __shared__ __attribute__((aligned(16))) uint16_t h2s[kP * 64];
  __shared__ __attribute__((aligned(16))) uint16_t h2t[64 * kP];
  __shared__ __attribute__((aligned(16))) uint16_t dys[kP * 32];   // [sample][out (pad 32)]
  __shared__ __attribute__((aligned(16))) uint16_t dyt[16 * kP];   // [out (pad 16)][sample]
  __shared__ __attribute__((aligned(16))) uint16_t dz2s[kP * 64];  // [sample][hid]
  __shared__ __attribute__((aligned(16))) uint16_t dz2t[64 * kP];  // [hid][sample]
  __shared__ __attribute__((aligned(16))) uint16_t dz1t[64 * kP];  // [hid][sample]
  __shared__ float s_db[4][2 * 64 + 16];
  __shared__ float s_loss[4];

  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int col = lane & 15, rq = (lane >> 4) * 4;
  const uint16_t* pk = A.P.packed;

  // zero the padded output columns once (rows 3..31 of dY stay zero)
  for (int i = t; i < kP * 32; i += 256) dys[i] = 0;
  for (int i = t; i < 16 * kP; i += 256) dyt[i] = 0;

  // gradient tiles owned by this wave: 4 (W3) + 16 (W2) + 12 (W1, in < 48) = 32
  floatx4 g[8];
  for (int q = 0; q < 8; ++q) g[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  float db1[16], db2[16], db3[3] = {0.f, 0.f, 0.f};
  for (int q = 0; q < 16; ++q) db1[q] = db2[q] = 0.f;
  float loss = 0.f;

  const uint32_t nchunks = (A.n + kP - 1) / kP;
  for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const uint32_t base = ch * kP;
    __syncthreads();
    // stage X in both layouts
    for (int i = t; i < kP * 8; i += 256) {
      const int s = i >> 3, q = i & 7;
      const uint32_t k = base + s;
      const uint4 v = k < A.n ? reinterpret_cast<const uint4*>(A.X + (size_t)k * 64)[q] : make_uint4(0, 0, 0, 0);
      *reinterpret_cast<uint4*>(xs + s * 64 + q * 8) = v;
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      for (int e = 0; e < 8; ++e) xt[(q * 8 + e) * kP + s] = (uint16_t)(w[e >> 1] >> ((e & 1) * 16));
    }
    __syncthreads();
    // forward layers 1 and 2 on the wave's 16-sample column tile
    floatx4 h1[4], h2[4];
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(pk + kPW1 + mt * 16 * 64, 64, xs + wave * 16 * 64, 64, 64, acc);
      for (int i = 0; i < 4; ++i) acc[i] = fmaxf(acc[i] + A.P.b1[mt * 16 + rq + i], 0.f);
      h1[mt] = acc;
      uint16_t o[4];
      for (int i = 0; i < 4; ++i) o[i] = bf16_bits(acc[i]);
      const int s = wave * 16 + col;
      *reinterpret_cast<uint2*>(h1s + s * 64 + mt * 16 + rq) =
          make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
      for (int i = 0; i < 4; ++i) h1t[(mt * 16 + rq + i) * kP + s] = o[i];
    }
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(pk + kPW2 + mt * 16 * 64, 64, h1s + wave * 16 * 64, 64, 64, acc);
      for (int i = 0; i < 4; ++i) acc[i] = fmaxf(acc[i] + A.P.b2[mt * 16 + rq + i], 0.f);
      h2[mt] = acc;
      uint16_t o[4];
      for (int i = 0; i < 4; ++i) o[i] = bf16_bits(acc[i]);
      const int s = wave * 16 + col;
      *reinterpret_cast<uint2*>(h2s + s * 64 + mt * 16 + rq) =
          make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
      for (int i = 0; i < 4; ++i) h2t[(mt * 16 + rq + i) * kP + s] = o[i];
    }
    __syncthreads();
    // layer 3 and the output gradient
    {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(pk + kPW3, 64, h2s + wave * 16 * 64, 64, 64, acc);
      const int s = wave * 16 + col;
      const uint32_t k = base + s;
      if (lane < 16) {  // rows 0..3 of the output tile: outputs 0..2
        for (int o = 0; o < kOut; ++o) {
          float d = 0.f;
          if (k < A.n) {
            const float yv = acc[o] + A.P.b3[o];
            const float a = A.Alb[(size_t)k * 3 + o];
            const float e = a * (yv - A.Y[(size_t)k * 3 + o]);
            loss += e * e;
            d = A.scale * a * e;  // d(loss)/dy
          }
          db3[o] += d;
          const uint16_t b = bf16_bits(d);
          dys[s * 32 + o] = b;
          dyt[o * kP + s] = b;
        }
      }
    }
    __syncthreads();
    // dH2 = W3^T dY -> dZ2 = dH2 * [h2 > 0]
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(pk + kPW3T + mt * 16 * 32, 32, dys + wave * 16 * 32, 32, 32, acc);
      uint16_t o[4];
      for (int i = 0; i < 4; ++i) {
        const float dz = h2[mt][i] > 0.f ? acc[i] : 0.f;
        db2[mt * 4 + i] += dz;
        o[i] = bf16_bits(dz);
      }
      const int s = wave * 16 + col;
      *reinterpret_cast<uint2*>(dz2s + s * 64 + mt * 16 + rq) =
          make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
      for (int i = 0; i < 4; ++i) dz2t[(mt * 16 + rq + i) * kP + s] = o[i];
    }
    __syncthreads();
    // dH1 = W2^T dZ2 -> dZ1 = dH1 * [h1 > 0]
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(pk + kPW2T + mt * 16 * 64, 64, dz2s + wave * 16 * 64, 64, 64, acc);
      const int s = wave * 16 + col;
      for (int i = 0; i < 4; ++i) {
        const float dz = h1[mt][i] > 0.f ? acc[i] : 0.f;
        db1[mt * 4 + i] += dz;
        dz1t[(mt * 16 + rq + i) * kP + s] = bf16_bits(dz);
      }
    }
    __syncthreads();
    // weight gradients over the chunk's 64 samples (K = samples)
    for (int q = 0; q < 8; ++q) {
      const int tile = wave * 8 + q;  // 0..31
      if (tile < 4) {                 // dW3 [16 x 64]: n tile = tile
        g[q] = mfma_tile(dyt, kP, h2t + tile * 16 * kP, kP, kP, g[q]);
      } else if (tile < 20) {         // dW2 [64 x 64]
        const int tt = tile - 4, mt = tt >> 2, nt = tt & 3;
        g[q] = mfma_tile(dz2t + mt * 16 * kP, kP, h1t + nt * 16 * kP, kP, kP, g[q]);
      } else {                        // dW1 [64 x 48]
        const int tt = tile - 20, mt = tt / 3, nt = tt % 3;
        g[q] = mfma_tile(dz1t + mt * 16 * kP, kP, xt + nt * 16 * kP, kP, kP, g[q]);
      }
    }
  }
  // ---- write this workgroup's partial gradients --------------------------------------
  float* out = A.partial + (size_t)blockIdx.x * (kParams + 1);
  for (int q = 0; q < 8; ++q) {
    const int tile = wave * 8 + q;
    for (int i = 0; i < 4; ++i) {
      const int m = rq + i;
      if (tile < 4) {
        if (m < kOut) out[kOffW3 + m * 64 + tile * 16 + col] = g[q][i];
      } else if (tile < 20) {
        const int tt = tile - 4, mt = tt >> 2, nt = tt & 3;
        out[kOffW2 + (mt * 16 + m) * 64 + nt * 16 + col] = g[q][i];
      } else {
        const int tt = tile - 20, mt = tt / 3, nt = tt % 3;
        const int in = nt * 16 + col;
        if (in < kIn) out[kOffW1 + (mt * 16 + m) * kIn + in] = g[q][i];
      }
    }
  }
  // bias gradients: sum the 16 sample lanes of each row group, then the waves
  for (int q = 0; q < 16; ++q) {
    float a = db1[q], b = db2[q];
    for (int o = 1; o < 16; o <<= 1) {
      a += __shfl_xor(a, o, 64);
      b += __shfl_xor(b, o, 64);
    }
    db1[q] = a;
    db2[q] = b;
  }
  if (col == 0) {
    for (int mt = 0; mt < 4; ++mt)
      for (int i = 0; i < 4; ++i) {
        s_db[wave][mt * 16 + rq + i] = db1[mt * 4 + i];
        s_db[wave][64 + mt * 16 + rq + i] = db2[mt * 4 + i];
      }
  }
  {
    float l = loss;
    for (int o = 1; o < 64; o <<= 1) l += __shfl_xor(l, o, 64);
    if (lane == 0) s_loss[wave] = l;
  }
  for (int o = 0; o < kOut; ++o) {
    float a = lane < 16 ? db3[o] : 0.f;
    for (int q = 1; q < 16; q <<= 1) a += __shfl_xor(a, q, 64);
    if (lane == 0) s_db[wave][128 + o] = a;
  }
  __syncthreads();
  if (t < 64) {
    out[kOffB1 + t] = ((s_db[0][t] + s_db[1][t]) + s_db[2][t]) + s_db[3][t];
    out[kOffB2 + t] = ((s_db[0][64 + t] + s_db[1][64 + t]) + s_db[2][64 + t]) + s_db[3][64 + t];
  }
  if (t < kOut) out[kOffB3 + t] = ((s_db[0][128 + t] + s_db[1][128 + t]) + s_db[2][128 + t]) + s_db[3][128 + t];
  if (t == 0) out[kParams] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// Fixed-order reduction of the partials + Adam (FP32 master weights).
// With lr == 0 it only reduces (gradient and loss read-back).
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ params, float* __restrict__ m1,
                                                       float* __restrict__ m2, const float* __restrict__ partial,
                                                       int nblk, float lr, float beta1, float beta2, float eps,
                                                       float bc1, float bc2, float* __restrict__ grad_out,
                                                       float* __restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  float g = 0.f;
  for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * (kParams + 1) + p];
  if (p == kParams) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (grad_out) grad_out[p] = g;
  if (lr > 0.f) {
    const float a = beta1 * m1[p] + (1.f - beta1) * g;
    const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  }
}

// BF16 operand copies of the master weights.
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ params, uint16_t* __restrict__ pk,
                                                       float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 64 * 64) {
    const int r = i >> 6, c = i & 63;
    pk[kPW1 + i] = c < kIn ? bf16_bits(params[kOffW1 + r * kIn + c]) : (uint16_t)0;
    pk[kPW2 + i] = bf16_bits(params[kOffW2 + r * 64 + c]);
    pk[kPW2T + i] = bf16_bits(params[kOffW2 + c * 64 + r]);  // [in r][out c]
  }
  if (i < 16 * 64) {
    const int r = i >> 6, c = i & 63;
    pk[kPW3 + i] = r < kOut ? bf16_bits(params[kOffW3 + r * 64 + c]) : (uint16_t)0;
  }
  if (i < 64 * 32) {
    const int r = i >> 5, c = i & 31;  // [hid r][out c]
    pk[kPW3T + i] = c < kOut ? bf16_bits(params[kOffW3 + c * 64 + r]) : (uint16_t)0;
  }
  if (i < 64) {
    bias[i] = params[kOffB1 + i];
    bias[64 + i] = params[kOffB2 + i];
  }
  if (i < 16) bias[128 + i] = i < kOut ? params[kOffB3 + i] : 0.f;
}

}  // namespace vct

namespace vct {

// ---- per-tile check and refinement -------------------------------------------------
// Every 32x32 tile is checked at 16 probe pixels traced at full resolution.
// Stage 1 probes sit at offsets 3, 11, 19, 27 (the odd-odd pixels farthest
// from the quarter-resolution samples).  By default the probes estimate each
// tile's error and the tiles with the largest estimates are traced in full
// until the estimated error of the rest fits 30% of the frame's budget for
// the bound (tile_budget_kernel).  With the half-resolution stage on, a tile
// whose mean squared error over its covered probes exceeds the bound has its
// odd-odd pixels (a 16x16 grid) traced and interpolated with the same surface
// guidance, is checked again at even-even probes (offsets 4, 12, 20, 28, the
// pixels farthest from that grid), and is traced in full if it still fails.
// Probe values are written into the image; a probe on a pixel that the
// interpolation left for tracing (w < 0) counts as exact.
constexpr int kProbes = 16;  // per tile: one or two 4x4 grids
static_assert(kProbes == 16 || kProbes == 32, "probe grids");

// One workgroup.  Checks every tile (in_list null) or the listed tiles,
// lists the failing ones in tile order and sets flags[t] to fail_flag or
// pass_flag.  Probe k of the i-th checked tile is probe_val / probe_pix [16 i + k].
__global__ __launch_bounds__(1024) void tile_select_kernel(const float4* __restrict__ probe_val,
                                                           const uint32_t* __restrict__ probe_pix,
                                                           const float4* __restrict__ gpos, float4* image,
                                                           const uint32_t* __restrict__ in_list, uint32_t ntiles,
                                                           float max_mse, uint32_t* __restrict__ out,
                                                           uint32_t* __restrict__ flags, uint32_t pass_flag,
                                                           uint32_t fail_flag) {
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_base;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_base = 0;
  __syncthreads();
  const uint32_t n = in_list ? in_list[0] : ntiles;
  uint32_t* list = out + 1;
  for (uint32_t i0 = 0; i0 < n; i0 += 1024) {
    const uint32_t i = i0 + tid;
    bool flag = false;
    uint32_t t = 0;
    if (i < n) {
      t = in_list ? in_list[1 + i] : i;
      float sse = 0.0f;
      int nc = 0;
      for (int k = 0; k < kProbes; ++k) {
        const uint32_t q = i * kProbes + k;
        const uint32_t p = probe_pix[q];
        if (gpos[p].w == 0.0f) continue;  // background
        const float4 v = probe_val[q], m = image[p];
        if (m.w < 0.0f) continue;  // left for tracing: exact
        const float dx = v.x - m.x, dy = v.y - m.y, dz = v.z - m.z;
        sse += dx * dx + dy * dy + dz * dz;
        ++nc;
      }
      flag = nc > 0 && sse > max_mse * 3.0f * (float)nc;
      for (int k = 0; k < kProbes; ++k) image[probe_pix[i * kProbes + k]] = probe_val[i * kProbes + k];
      flags[t] = flag ? fail_flag : pass_flag;
    }
    const unsigned long long m = __ballot(flag);
    if (lane == 0) s_wsum[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = s_base;
    for (int w = 0; w < wave; ++w) off += s_wsum[w];
    if (flag) list[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = t;
    __syncthreads();
    if (tid == 0) {
      uint32_t tot = 0;
      for (int w = 0; w < 16; ++w) tot += s_wsum[w];
      s_base += tot;
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = s_base;
}

// Stage-1 estimate: est[t] = sum over the tile's covered probes of the
// squared error (3 channels) of the learned value, ncov[t] = covered probes;
// with write_probes the probe values then go into the image.  A probe on a
// pixel the interpolation left for tracing (w < 0) counts as exact.
__global__ __launch_bounds__(256) void tile_estimate_kernel(const float4* __restrict__ probe_val,
                                                            const uint32_t* __restrict__ probe_pix,
                                                            const float4* __restrict__ gpos, float4* image,
                                                            uint32_t ntiles, float* __restrict__ est,
                                                            uint32_t* __restrict__ ncov, int write_probes) {
  // one lane per probe, kProbes lanes per tile (a power of two dividing 64);
  // the tile's sums are reduced across its lanes in a fixed order
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t t = g / kProbes;
  float e = 0.0f;
  uint32_t c = 0;
  if (t < ntiles) {
    const uint32_t p = probe_pix[g];
    const float4 v = probe_val[g], m = image[p];
    if (gpos[p].w != 0.0f) {  // covered
      c = 1;
      if (m.w >= 0.0f) {  // w < 0: left for tracing, exact
        const float dx = v.x - m.x, dy = v.y - m.y, dz = v.z - m.z;
        e = dx * dx + dy * dy + dz * dz;
      }
    }
  }
  for (int o = kProbes / 2; o; o >>= 1) {
    e += __shfl_xor(e, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if (t < ntiles) {
    if (write_probes) image[probe_pix[g]] = probe_val[g];
    if (g % kProbes == 0) {
      est[t] = e;
      ncov[t] = c;
    }
  }
}

// One workgroup: the tiles to trace in full, largest estimated error first,
// until the error left in the others fits the frame's budget,
//   sum_{not picked} est <= 3 * max_mse * share * sum ncov
// (3 channels; each covered probe stands for the same share of the tile's
// covered pixels).  The rest of the budget is margin for the estimate: 16
// probes per tile estimate a tile's error loosely, and the sum over n tiles
// with covered probes errs relatively by ~1/sqrt(n), so
//   share = clamp(0.55 - 1.7 / sqrt(n), 0.25, 0.55).
// The constant is set by the smallest frame measured: at ~40 covered tiles a
// share of 0.5 ended at 38.6 dB and 0.3 at 40.9 dB for a 40 dB bound.  The estimates are binned by their float exponent and two
// mantissa bits (4 bins per octave) and summed in 2^-40 fixed point, so the
// pick is deterministic; every tile of the threshold bin is picked.
// flags[t] = 1 picked, 0 not; out = [count, picked tiles...].
constexpr int kBudgetBins = 256;
constexpr float kShareMax = 0.55f, kShareMin = 0.25f, kShareSpread = 1.7f;

__device__ __forceinline__ int budget_bin(float e) {
  if (!(e > 0.0f)) return -1;  // no error: never picked
  const int b = (int)(__float_as_uint(e) >> 21) - 300;  // 2^-52 .. 2^12
  return b < 0 ? 0 : (b >= kBudgetBins ? kBudgetBins - 1 : b);
}

__global__ __launch_bounds__(1024) void tile_budget_kernel(const float* __restrict__ est,
                                                           const uint32_t* __restrict__ ncov, uint32_t ntiles,
                                                           float max_mse, uint32_t* __restrict__ out,
                                                           uint32_t* __restrict__ flags) {
  __shared__ unsigned long long s_mass[kBudgetBins];
  __shared__ unsigned long long s_total;
  __shared__ unsigned int s_cov, s_count, s_tiles;
  __shared__ int s_thr;
  const int tid = threadIdx.x;
  const float kFix = 1099511627776.0f;  // 2^40
  for (int b = tid; b < kBudgetBins; b += 1024) s_mass[b] = 0ull;
  if (tid == 0) {
    s_total = 0ull;
    s_cov = 0u;
    s_count = 0u;
    s_tiles = 0u;
  }
  __syncthreads();
  {
    // per-bin masses by LDS atomics; the totals summed per wave first (exact
    // integers, so the order does not matter)
    unsigned long long tot = 0ull;
    uint32_t cov = 0u, tiles = 0u;
    for (uint32_t t = tid; t < ntiles; t += 1024) {
      const float e = est[t];
      const int b = budget_bin(e);
      cov += ncov[t];
      tiles += ncov[t] ? 1u : 0u;
      if (b >= 0) {
        const unsigned long long q = (unsigned long long)(e * kFix);
        atomicAdd(&s_mass[b], q);
        tot += q;
      }
    }
    for (int o = 32; o; o >>= 1) {
      tot += __shfl_xor(tot, o, 64);
      cov += __shfl_xor(cov, o, 64);
      tiles += __shfl_xor(tiles, o, 64);
    }
    if ((tid & 63) == 0) {
      atomicAdd(&s_total, tot);
      atomicAdd(&s_cov, cov);
      atomicAdd(&s_tiles, tiles);
    }
  }
  __syncthreads();
  if (tid == 0) {
    const double share = fmin(fmax((double)kShareMax - (double)kShareSpread / sqrt(fmax((double)s_tiles, 1.0)),
                                   (double)kShareMin), (double)kShareMax);
    const double budget = 3.0 * (double)max_mse * share * (double)s_cov * (double)kFix;
    double left = (double)s_total;
    int thr = kBudgetBins;  // pick bins >= thr
    while (thr > 0 && left > budget) {
      --thr;
      left -= (double)s_mass[thr];
    }
    s_thr = thr;
  }
  __syncthreads();
  const int thr = s_thr;
  const int lane = tid & 63;
  for (uint32_t t0 = 0; t0 < ntiles; t0 += 1024) {
    const uint32_t t = t0 + tid;
    const bool pick = t < ntiles && budget_bin(est[t]) >= thr;
    if (t < ntiles) flags[t] = pick ? 1u : 0u;
    const unsigned long long m = __ballot(pick);
    uint32_t base = 0;
    if (lane == 0 && m) base = atomicAdd(&s_count, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    if (pick) out[1 + base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = t;
  }
  __syncthreads();
  if (tid == 0) out[0] = s_count;
}

// Stage-2 probe pixels of the listed tiles: pix[0] = count, then 16 per tile.
__global__ void stage2_probes_kernel(const uint32_t* __restrict__ list, int tiles_x, int W, int H,
                                     uint32_t* __restrict__ pix) {
  const uint32_t n = list[0];
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n * kProbes; q += gridDim.x * blockDim.x) {
    const uint32_t t = list[1 + q / kProbes], k = q % kProbes;
    const int o = (k >> 4) ? 0 : 4;  // 4x4 grids at offsets 4, then 0 (even-even pixels)
    const int x = min((int)(t % (uint32_t)tiles_x) * kFbTile + 8 * (int)(k & 3u) + o, W - 1);
    const int y = min((int)(t / (uint32_t)tiles_x) * kFbTile + 8 * (int)((k >> 2) & 3u) + o, H - 1);
    pix[1 + q] = (uint32_t)(y * W + x);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) pix[0] = n * kProbes;
}

// Half-resolution interpolation of the listed tiles (one workgroup per tile):
// the 3x3 odd-odd grid points around the pixel's nearest one, weighted by a
// tent of radius 1.25 grid points (2.5 pixels; wider tents blur the voxel-scale
// detail of the indirect light) and by the same-surface test of the
// quarter-resolution stage; grid points of tiles without half-resolution
// data do not count.  Grid points keep their traced value; pixels without a
// same-surface grid point are listed in trace_list and marked w = -1.
__global__ __launch_bounds__(256) void half_upsample_kernel(const uint32_t* __restrict__ list,
                                                            const uint32_t* __restrict__ flags, int tiles_x,
                                                            const float4* __restrict__ gpos,
                                                            const float4* __restrict__ gnrm,
                                                            const DevMaterial* __restrict__ mats, uint32_t nmat,
                                                            const float4* __restrict__ half, int hw, int hh, int W,
                                                            int H, float4* __restrict__ image,
                                                            uint32_t* __restrict__ trace_list) {
  const uint32_t k = blockIdx.x;
  if (k >= list[0]) return;
  const uint32_t t = list[1 + k];
  const int x0 = (int)(t % (uint32_t)tiles_x) * kFbTile, y0 = (int)(t / (uint32_t)tiles_x) * kFbTile;
  const int lane = threadIdx.x & 63;
  for (int q = threadIdx.x; q < kFbTile * kFbTile; q += 256) {
    const int px = x0 + (q & 31), py = y0 + (q >> 5);
    const bool inside = px < W && py < H;
    const size_t idx = inside ? (size_t)py * W + px : 0;
    const float4 gp = gpos[idx];
    int st = inside ? (gp.w == 0.0f ? kBackground : kFeatures) : -1;
    float4 o = make_float4(0.15f, 0.25f, 0.25f, 1.0f);  // clear colour, renderer.cpp:398
    if (st == kFeatures) {
      const float4 gn = gnrm[idx];
      const uint32_t mbits = __builtin_bit_cast(uint32_t, gn.w);
      const DevMaterial& m = mats[mbits < nmat ? mbits : 0u];
      const float nl = sqrtf(gn.x * gn.x + gn.y * gn.y + gn.z * gn.z);
      const float inl = nl > 0.0f ? 1.0f / nl : 0.0f;
      const float nx = gn.x * inl, ny = gn.y * inl, nz = gn.z * inl;
      const float u = (float)(px - 1) * 0.5f, v = (float)(py - 1) * 0.5f;
      float3 acc = make_float3(0.f, 0.f, 0.f);
      float wsum = 0.0f;
      if ((px & 1) && (py & 1)) {  // a grid point: its traced value
        const float4 h = half[(size_t)(py >> 1) * hw + (px >> 1)];
        acc = make_float3(h.x, h.y, h.z);
        wsum = 1.0f;
      } else {
        const int ic = (int)floorf(u + 0.5f), jc = (int)floorf(v + 0.5f);
        for (int kk = 0; kk < 9; ++kk) {
          const int i = ic + (kk % 3) - 1, j = jc + (kk / 3) - 1;
          if (i < 0 || i >= hw || j < 0 || j >= hh) continue;
          const int sx = 2 * i + 1, sy = 2 * j + 1;
          if (sx >= W || sy >= H || flags[(sy >> 5) * tiles_x + (sx >> 5)] == 0u) continue;
          const size_t sidx = (size_t)sy * W + sx;
          const float4 ln = gnrm[sidx];
          const float4 hv = half[(size_t)j * hw + i];
          if (__builtin_bit_cast(uint32_t, ln.w) != mbits || !(hv.w > 0.0f)) continue;
          const float d = nx * ln.x + ny * ln.y + nz * ln.z;
          const float ll = sqrtf(ln.x * ln.x + ln.y * ln.y + ln.z * ln.z);
          if (!(d > 0.95f * ll) || !(fabsf(hv.w - gp.w) < 0.05f * gp.w)) continue;
          const float wk =
              fmaxf(0.0f, 1.0f - 0.8f * fabsf(u - (float)i)) * fmaxf(0.0f, 1.0f - 0.8f * fabsf(v - (float)j));
          acc = make_float3(acc.x + wk * hv.x, acc.y + wk * hv.y, acc.z + wk * hv.z);
          wsum += wk;
        }
      }
      if (wsum > 0.0f) {
        const float iw = 1.0f / wsum;
        o = make_float4(m.diffuse[0] * acc.x * iw, m.diffuse[1] * acc.y * iw, m.diffuse[2] * acc.z * iw, 1.0f);
      } else {
        st = kTraceIt;
        o = make_float4(0.f, 0.f, 0.f, -1.0f);
      }
    }
    const unsigned long long bm = __ballot(st == kTraceIt);
    uint32_t b0 = 0;
    if (lane == 0 && bm) b0 = atomicAdd(trace_list, (uint32_t)__popcll(bm));
    b0 = __shfl(b0, 0, 64);
    if (st == kTraceIt) trace_list[1 + b0 + (uint32_t)__popcll(bm & ((1ull << lane) - 1ull))] = (uint32_t)idx;
    if (st >= 0) image[idx] = o;
  }
}

}  // namespace vct

// ============================================================================
// host side
// ============================================================================
using namespace vct;

namespace {

__global__ void lowres_gather_kernel(const float4* __restrict__ gpos, const float4* __restrict__ gnrm, int W, int H,
                                     float4* __restrict__ lpos, float4* __restrict__ lnrm, int wl, int hl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= wl * hl) return;
  const int x = i % wl, y = i / wl;
  const int sx = min(4 * x + 1, W - 1), sy = min(4 * y + 1, H - 1);
  lpos[i] = gpos[(size_t)sy * W + sx];
  lnrm[i] = gnrm[(size_t)sy * W + sx];
}

struct HostRng {
  uint64_t s;
  explicit HostRng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int mlp_ensure(vct_context* ctx) {
  int rc;
  if ((rc = vct_ensure(ctx->mlp_params, kParams * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->mlp_m1, kParams * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->mlp_m2, kParams * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->mlp_packed, kPacked * 2, ctx))) return rc;
  if ((rc = vct_ensure(ctx->mlp_bias, 144 * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->mlp_grad, (kParams + 1) * 4, ctx))) return rc;
  return VCT_OK;
}

int mlp_pack(vct_context* ctx) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(16), dim3(256), 0, ctx->stream, (const float*)ctx->mlp_params.p,
                     (uint16_t*)ctx->mlp_packed.p, (float*)ctx->mlp_bias.p);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

MlpDev mlp_dev(vct_context* ctx) {
  MlpDev P;
  P.packed = (const uint16_t*)ctx->mlp_packed.p;
  P.b1 = (const float*)ctx->mlp_bias.p;
  P.b2 = P.b1 + 64;
  P.b3 = P.b1 + 128;
  return P;
}

int need_mlp(vct_context* ctx) {
  if (!ctx->mlp_ready) return vct_set_error(VCT_ERR_STATE, "upsampler parameters not set (vct_mlp_init first)");
  return VCT_OK;
}

// Training step over X/A/Y rows [off, off + n) already on the device.
int mlp_train_on(vct_context* ctx, const uint16_t* X, const float* Alb, const float* Y, uint32_t n, float lr,
                 float* loss_host, float* grad_host) {
  const uint32_t nchunks = (n + kP - 1) / kP;
  const int G = (int)std::min<uint32_t>(nchunks, 1024u);
  int rc;
  if ((rc = vct_ensure(ctx->mlp_partial, (size_t)G * (kParams + 1) * 4, ctx))) return rc;
  TrainArgs A;
  A.P = mlp_dev(ctx);
  A.X = X;
  A.Alb = Alb;
  A.Y = Y;
  A.n = n;
  A.scale = 2.0f / (3.0f * (float)n);
  A.partial = (float*)ctx->mlp_partial.p;
  hipLaunchKernelGGL(mlp_train_kernel, dim3(G), dim3(256), 0, ctx->stream, A);
  VCT_HIP(hipGetLastError());
  float bc1 = 1.f, bc2 = 1.f;
  if (lr > 0.f) {
    ctx->mlp_step += 1;
    bc1 = 1.0f - powf(0.9f, (float)ctx->mlp_step);
    bc2 = 1.0f - powf(0.999f, (float)ctx->mlp_step);
  }
  float* gout = (float*)ctx->mlp_grad.p;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 1 + 255) / 256), dim3(256), 0, ctx->stream,
                     (float*)ctx->mlp_params.p, (float*)ctx->mlp_m1.p, (float*)ctx->mlp_m2.p,
                     (const float*)ctx->mlp_partial.p, G, lr, 0.9f, 0.999f, 1e-8f, bc1, bc2, gout, gout + kParams);
  VCT_HIP(hipGetLastError());
  if (lr > 0.f) {
    if ((rc = mlp_pack(ctx))) return rc;
  }
  if (loss_host || grad_host) {
    std::vector<float> g(kParams + 1);
    VCT_HIP(hipMemcpyAsync(g.data(), gout, (kParams + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
    if (grad_host) memcpy(grad_host, g.data(), kParams * 4);
    if (loss_host) *loss_host = g[kParams] / (3.0f * (float)n);
  }
  return VCT_OK;
}

int upload_features(vct_context* ctx, const float* features, uint32_t n, VctBuf& buf) {
  std::vector<uint16_t> h((size_t)n * 64, 0);
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < kIn; ++c) h[i * 64 + c] = f32_to_bf16(features[i * kIn + c]);
  int rc;
  if ((rc = vct_ensure(buf, h.size() * 2 + 16, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(buf.p, h.data(), h.size() * 2, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

FeatArgs feat_args(vct_context* ctx) {
  FeatArgs F;
  F.gpos = (const float4*)ctx->g_pos.p;
  F.gnrm = (const float4*)ctx->g_nrm.p;
  F.mats = (const DevMaterial*)ctx->d_materials.p;
  F.nmat = (uint32_t)ctx->materials.size();
  F.low = (const float4*)ctx->low_img.p;
  F.lnrm = (const float4*)ctx->low_gnrm.p;
  F.W = ctx->width;
  F.H = ctx->height;
  F.wl = ctx->width / 4;
  F.hl = ctx->height / 4;
  return F;
}

// The quarter-resolution buffers (pixel (4i+1, 4j+1) of every 4x4 block).
int lowres_buffers(vct_context* ctx) {
  const int W = ctx->width, H = ctx->height;
  if (W % 4 || H % 4) return vct_set_error(VCT_ERR_STATE, "the learned path needs a viewport divisible by 4");
  const size_t nl = (size_t)(W / 4) * (H / 4);
  int rc;
  if ((rc = vct_ensure(ctx->low_gpos, nl * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->low_gnrm, nl * 16, ctx))) return rc;
  return vct_ensure(ctx->low_img, nl * 16, ctx);
}

int launch_lowres_gather(vct_context* ctx) {
  const int W = ctx->width, H = ctx->height, wl = W / 4, hl = H / 4;
  const size_t nl = (size_t)wl * hl;
  hipLaunchKernelGGL(lowres_gather_kernel, dim3((uint32_t)((nl + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p, W, H, (float4*)ctx->low_gpos.p,
                     (float4*)ctx->low_gnrm.p, wl, hl);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// Quarter-resolution G-buffer and the albedo-demodulated indirect light
// traced there (w = linear depth); `gathered`: the frame front already
// gathered the G-buffer (on the G-buffer's stream, joined before the mip's
// end), so only the trace is left.
int trace_lowres(vct_context* ctx, bool gathered = false) {
  int rc;
  if ((rc = lowres_buffers(ctx))) return rc;
  if (!gathered && (rc = launch_lowres_gather(ctx))) return rc;
  return vct_trace_into(ctx, (const float4*)ctx->low_gpos.p, (const float4*)ctx->low_gnrm.p,
                        (float4*)ctx->low_img.p, ctx->width / 4, ctx->height / 4, 1, nullptr);
}

__global__ void zero_count_kernel(uint32_t* p) { *p = 0u; }

// Upsampling, then the full-resolution trace of the pixels it listed.
int infer_full(vct_context* ctx, const uint32_t* skip_tile = nullptr) {
  const size_t npix = (size_t)ctx->width * ctx->height;
  int rc;
  if ((rc = vct_ensure(ctx->fb_pix, (npix + 1) * 4, ctx))) return rc;
  hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, ctx->stream, (uint32_t*)ctx->fb_pix.p);
  InferArgs A;
  memset(&A, 0, sizeof(A));
  A.F = feat_args(ctx);
  A.P = mlp_dev(ctx);
  A.image = (float4*)ctx->image.p;
  A.trace_list = (uint32_t*)ctx->fb_pix.p;
  A.skip_tile = skip_tile;
  A.tiles_x = (ctx->width + kFbTile - 1) / kFbTile;
  dim3 grid((ctx->width + 15) / 16, (ctx->height + 15) / 16);
  hipLaunchKernelGGL(mlp_infer_kernel, grid, dim3(256), 0, ctx->stream, A);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

int trace_listed_pixels(vct_context* ctx) {
  TraceSelect sel;
  memset(&sel, 0, sizeof(sel));
  sel.pix_list = (const uint32_t*)ctx->fb_pix.p + 1;
  sel.pix_count = (const uint32_t*)ctx->fb_pix.p;
  sel.npix_list = (uint32_t)((size_t)ctx->width * ctx->height);
  sel.pix_scatter = 1;
  return vct_trace_select(ctx, (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p, (float4*)ctx->image.p,
                          ctx->width, ctx->height, 0, nullptr, sel);
}

// Stage-1 probes of every tile, traced at full resolution into probe_val.
// They need only the G-buffer and the volume, so the learned frame traces
// them on the second stream beside the quarter-resolution trace and the MLP.
int refine_probes(vct_context* ctx) {
  const int W = ctx->width, H = ctx->height;
  const int tx = (W + kFbTile - 1) / kFbTile, ty = (H + kFbTile - 1) / kFbTile;
  const uint32_t ntiles = (uint32_t)(tx * ty);
  int rc;
  if ((rc = vct_ensure(ctx->probe_val, (size_t)ntiles * kProbes * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->fb_list, (size_t)(ntiles + 1) * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->fb_flag, (size_t)ntiles * 4, ctx))) return rc;
  if (ctx->probe_w != W || ctx->probe_h != H || !ctx->probe_pix.p) {
    std::vector<uint32_t> pix((size_t)ntiles * kProbes);
    for (int j = 0; j < ty; ++j)
      for (int i = 0; i < tx; ++i)
        for (int g = 0; g < kProbes / 16; ++g)  // 4x4 grids at offsets 3, then 7
          for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) {
              const int o = g ? 7 : 3;
              const int x = std::min(i * kFbTile + 8 * a + o, W - 1), y = std::min(j * kFbTile + 8 * b + o, H - 1);
              pix[((size_t)(j * tx + i)) * kProbes + g * 16 + b * 4 + a] = (uint32_t)(y * W + x);
            }
    if ((rc = vct_ensure(ctx->probe_pix, pix.size() * 4, ctx))) return rc;
    VCT_HIP(hipMemcpyAsync(ctx->probe_pix.p, pix.data(), pix.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->probe_w = W;
    ctx->probe_h = H;
  }
  TraceSelect sel;
  memset(&sel, 0, sizeof(sel));
  sel.pix_list = (const uint32_t*)ctx->probe_pix.p;
  sel.npix_list = ntiles * kProbes;
  return vct_trace_select(ctx, (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p,
                          (float4*)ctx->probe_val.p, W, H, 0, nullptr, sel);
}

// Stage-1 selection of the frame budget: estimates from the probes and the
// learned values at the probe pixels, then the tiles to trace in full (flag 1,
// listed in fb_list).  write_probes puts the probe values into the image.
int budget_select(vct_context* ctx, bool write_probes) {
  const int W = ctx->width, H = ctx->height;
  const int tx = (W + kFbTile - 1) / kFbTile, ty = (H + kFbTile - 1) / kFbTile;
  const uint32_t ntiles = (uint32_t)(tx * ty);
  int rc;
  if ((rc = vct_ensure(ctx->fb_list, (size_t)(ntiles + 1) * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->fb_flag, (size_t)ntiles * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->fb_est, (size_t)ntiles * 8, ctx))) return rc;
  const float max_mse = powf(10.0f, -ctx->fb_psnr_db / 10.0f);
  float* est = (float*)ctx->fb_est.p;
  uint32_t* ncov = (uint32_t*)(est + ntiles);
  hipLaunchKernelGGL(tile_estimate_kernel, dim3((ntiles * kProbes + 255) / 256), dim3(256), 0, ctx->stream,
                     (const float4*)ctx->probe_val.p, (const uint32_t*)ctx->probe_pix.p, (const float4*)ctx->g_pos.p,
                     (float4*)ctx->image.p, ntiles, est, ncov, write_probes ? 1 : 0);
  hipLaunchKernelGGL(tile_budget_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const float*)est,
                     (const uint32_t*)ncov, ntiles, max_mse, (uint32_t*)ctx->fb_list.p, (uint32_t*)ctx->fb_flag.p);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

__global__ void probe_write_kernel(const float4* __restrict__ val, const uint32_t* __restrict__ pix, uint32_t n,
                                   float4* __restrict__ image) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) image[pix[k]] = val[k];
}

// The learned values at the probe pixels only (no trace list).
int infer_probes(vct_context* ctx) {
  const int tx = (ctx->width + kFbTile - 1) / kFbTile, ty = (ctx->height + kFbTile - 1) / kFbTile;
  const uint32_t n = (uint32_t)(tx * ty) * kProbes;
  InferArgs A;
  memset(&A, 0, sizeof(A));
  A.F = feat_args(ctx);
  A.P = mlp_dev(ctx);
  A.image = (float4*)ctx->image.p;
  A.pix_list = (const uint32_t*)ctx->probe_pix.p;
  A.npix_list = n;
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// Per-tile check and refinement of the learned frame, after refine_probes
// (see tile_select_kernel and tile_budget_kernel); `selected`: the frame
// budget's selection already ran (budget_select) and the probes are in the
// image.
int refine(vct_context* ctx, bool selected) {
  const int W = ctx->width, H = ctx->height;
  const int tx = (W + kFbTile - 1) / kFbTile, ty = (H + kFbTile - 1) / kFbTile;
  const uint32_t ntiles = (uint32_t)(tx * ty);
  const size_t npix = (size_t)W * H;
  const float4* gpos = (const float4*)ctx->g_pos.p;
  const float4* gnrm = (const float4*)ctx->g_nrm.p;
  float4* image = (float4*)ctx->image.p;
  int rc;
  if ((rc = vct_ensure(ctx->fb_list, (size_t)(ntiles + 1) * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->fb_flag, (size_t)ntiles * 4, ctx))) return rc;
  const float max_mse = powf(10.0f, -ctx->fb_psnr_db / 10.0f);
  uint32_t* flags = (uint32_t*)ctx->fb_flag.p;
  uint32_t* l1 = (uint32_t*)ctx->fb_list.p;
  // stage 1: picked tiles -> l1 (flag 1), others flag 0
  TraceSelect sel;
  if (ctx->fb_half) {  // per-tile bound, then the half-resolution stage
    hipLaunchKernelGGL(tile_select_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const float4*)ctx->probe_val.p,
                       (const uint32_t*)ctx->probe_pix.p, gpos, image, (const uint32_t*)nullptr, ntiles, max_mse, l1,
                       flags, 0u, 1u);
  } else if (!selected) {  // frame budget: largest estimated errors first
    if ((rc = budget_select(ctx, true))) return rc;
  }
  VCT_HIP(hipGetLastError());
  // pixels the quarter-resolution stage left, in tiles that passed: on the
  // second stream, beside the tiles traced in full (disjoint pixels)
  memset(&sel, 0, sizeof(sel));
  sel.pix_list = (const uint32_t*)ctx->fb_pix.p + 1;
  sel.pix_count = (const uint32_t*)ctx->fb_pix.p;
  sel.npix_list = (uint32_t)npix;
  sel.pix_scatter = 1;
  sel.tile_flag = flags;
  sel.flag_want = 0u;
  sel.tiles_x = tx;
  VCT_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
  VCT_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  std::swap(ctx->stream, ctx->side);
  rc = vct_trace_select(ctx, gpos, gnrm, image, W, H, 0, nullptr, sel);
  std::swap(ctx->stream, ctx->side);
  if (rc) return rc;
  VCT_HIP(hipEventRecord(ctx->ev_join, ctx->side));
  const uint32_t* full_list = l1;
  ctx->fb_half_ran = false;
  if (ctx->fb_half) {
    const int hw = (W + 1) / 2, hh = (H + 1) / 2;
    if ((rc = vct_ensure(ctx->half_img, (size_t)hw * hh * 16, ctx))) return rc;
    if ((rc = vct_ensure(ctx->fb_pix2, (npix + 1) * 4, ctx))) return rc;
    if ((rc = vct_ensure(ctx->fb_list2, (size_t)(ntiles + 1) * 4, ctx))) return rc;
    if ((rc = vct_ensure(ctx->probe_pix2, ((size_t)ntiles * kProbes + 1) * 4, ctx))) return rc;
    if ((rc = vct_ensure(ctx->probe_val2, (size_t)ntiles * kProbes * 16, ctx))) return rc;
    // odd-odd grid of the l1 tiles, demodulated (w = depth)
    memset(&sel, 0, sizeof(sel));
    sel.tile_list = l1 + 1;
    sel.tile_count = l1;
    sel.max_tiles = ntiles;
    sel.tiles_x = tx;
    sel.grid_stride = 2;
    sel.half_w = hw;
    if ((rc = vct_trace_select(ctx, gpos, gnrm, (float4*)ctx->half_img.p, W, H, 1, nullptr, sel))) return rc;
    uint32_t* p2 = (uint32_t*)ctx->fb_pix2.p;
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, ctx->stream, p2);
    hipLaunchKernelGGL(half_upsample_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, (const uint32_t*)l1,
                       (const uint32_t*)flags, tx, gpos, gnrm, (const DevMaterial*)ctx->d_materials.p,
                       (uint32_t)ctx->materials.size(), (const float4*)ctx->half_img.p, hw, hh, W, H, image, p2);
    VCT_HIP(hipGetLastError());
    // stage 2: even-even probes of the l1 tiles -> l2 (flag 2), others stay 1
    uint32_t* pp2 = (uint32_t*)ctx->probe_pix2.p;
    hipLaunchKernelGGL(stage2_probes_kernel, dim3(64), dim3(256), 0, ctx->stream, (const uint32_t*)l1, tx, W, H, pp2);
    VCT_HIP(hipGetLastError());
    memset(&sel, 0, sizeof(sel));
    sel.pix_list = pp2 + 1;
    sel.pix_count = pp2;
    sel.npix_list = ntiles * kProbes;
    if ((rc = vct_trace_select(ctx, gpos, gnrm, (float4*)ctx->probe_val2.p, W, H, 0, nullptr, sel))) return rc;
    uint32_t* l2 = (uint32_t*)ctx->fb_list2.p;
    hipLaunchKernelGGL(tile_select_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const float4*)ctx->probe_val2.p,
                       (const uint32_t*)(pp2 + 1), gpos, image, (const uint32_t*)l1, ntiles, max_mse, l2, flags, 1u,
                       2u);
    VCT_HIP(hipGetLastError());
    // pixels the half-resolution stage left, in tiles that passed stage 2
    memset(&sel, 0, sizeof(sel));
    sel.pix_list = p2 + 1;
    sel.pix_count = p2;
    sel.npix_list = (uint32_t)npix;
    sel.pix_scatter = 1;
    sel.tile_flag = flags;
    sel.flag_want = 1u;
    sel.tiles_x = tx;
    if ((rc = vct_trace_select(ctx, gpos, gnrm, image, W, H, 0, nullptr, sel))) return rc;
    full_list = l2;
    ctx->fb_half_ran = true;
  }
  // tiles traced in full
  memset(&sel, 0, sizeof(sel));
  sel.tile_list = full_list + 1;
  sel.tile_count = full_list;
  sel.max_tiles = ntiles;
  sel.tiles_x = tx;
  if ((rc = vct_trace_select(ctx, gpos, gnrm, image, W, H, 0, nullptr, sel))) return rc;
  VCT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  return VCT_OK;
}

// Minibatch k of a training run: rows perm[(k*batch + i) mod n] of the
// sample set, gathered into contiguous batch buffers.  Eight threads per row
// move its 128-byte BF16 feature row as 16-byte pieces; the first of them
// also moves albedo and target.
__global__ __launch_bounds__(256) void mlp_batch_gather_kernel(const uint32_t* __restrict__ perm, uint32_t n,
                                                               uint64_t first, uint32_t batch,
                                                               const uint16_t* __restrict__ X,
                                                               const float* __restrict__ A,
                                                               const float* __restrict__ Y, uint16_t* __restrict__ bx,
                                                               float* __restrict__ ba, float* __restrict__ by) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = t >> 3, c = t & 7;
  if (i >= batch) return;
  const uint32_t r = perm[(first + i) % n];
  reinterpret_cast<uint4*>(bx + (size_t)i * 64)[c] = reinterpret_cast<const uint4*>(X + (size_t)r * 64)[c];
  if (c == 0) {
    for (int q = 0; q < 3; ++q) {
      ba[(size_t)i * 3 + q] = A[(size_t)r * 3 + q];
      by[(size_t)i * 3 + q] = Y[(size_t)r * 3 + q];
    }
  }
}

// A uniformly random order of the n training rows (Fisher-Yates, the
// library's splitmix64 stream of `seed`), on the device in ctx->train_perm.
int shuffle_rows(vct_context* ctx, uint32_t n, uint64_t seed) {
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  HostRng rng(seed);
  for (uint32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.next() % ((uint64_t)i + 1)]);
  int rc;
  if ((rc = vct_ensure(ctx->train_perm, (size_t)n * 4, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->train_perm.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));  // `perm` is pageable and goes out of scope
  return VCT_OK;
}

}  // namespace

int vct_launch_lowres_gather(vct_context* ctx) { return launch_lowres_gather(ctx); }

#define MLP_CTX(ctx)                                                          \
  do {                                                                        \
    if (!(ctx)) return vct_set_error(VCT_ERR_BAD_HANDLE, "null context");     \
    VCT_HIP(hipSetDevice((ctx)->device));                                     \
  } while (0)

// ---- the forward pass that allocates nothing ----------------------------------------
// The BF16 feature rows of vct_mlp_forward_noalloc are staged in this fixed
// array: each GPU the library runs on holds its own copy until the process
// ends, at its own address.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in symbol): the device address comes
// from hipGetSymbolAddress, taken once per device when a context is created.
static __device__ __align__(256) uint16_t scratch[(size_t)VCT_MLP_NOALLOC_MAX_ROWS * 64];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint16_t* g_scratch_ptr[kScratchDevices];  // device address of `scratch` per device (null: not set up)
}  // namespace

// Sets up the device address of the staging array of the current device.
int vct_noalloc_setup(int device) {
  if (device < 0 || device >= kScratchDevices) return vct_set_error(VCT_ERR_LIMIT, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return VCT_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return vct_hip_check(e, "hipGetSymbolAddress(scratch)");
  if (!ptr) return vct_set_error(VCT_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch_ptr[device] = (uint16_t*)ptr;
  return VCT_OK;
}

// Feature rows (n x 34 floats) -> BF16 rows of 64 (columns 34..63 zero), the
// layout the forward kernel reads; one thread per (row, 8 columns).
__global__ __launch_bounds__(256) void mlp_stage_kernel(const float* __restrict__ features, uint32_t n,
                                                        uint16_t* __restrict__ X) {
  const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (g >= (uint64_t)n * 8u) return;
  const uint32_t row = (uint32_t)(g >> 3), q = (uint32_t)(g & 7u);
  uint32_t w[4];
  for (int e = 0; e < 8; e += 2) {
    const int c0 = (int)q * 8 + e, c1 = c0 + 1;
    const uint32_t lo = c0 < kIn ? f32_to_bf16(features[(size_t)row * kIn + c0]) : 0u;
    const uint32_t hi = c1 < kIn ? f32_to_bf16(features[(size_t)row * kIn + c1]) : 0u;
    w[e >> 1] = lo | (hi << 16);
  }
  reinterpret_cast<uint4*>(X + (size_t)row * 64)[q] = make_uint4(w[0], w[1], w[2], w[3]);
}

extern "C" {

int vct_mlp_init(vct_ctx ctx, uint64_t seed) {
  MLP_CTX(ctx);
  int rc;
  if ((rc = mlp_ensure(ctx))) return rc;
  std::vector<float> p(kParams, 0.0f);
  HostRng rng(seed);
  auto fill = [&](int off, int rows, int fan_in, double gain) {
    const double b = gain * std::sqrt(6.0 / fan_in);
    for (int i = 0; i < rows * fan_in; ++i) p[off + i] = (float)((2.0 * rng.uniform() - 1.0) * b);
  };
  fill(kOffW1, kHid, kIn, 1.0);
  fill(kOffW2, kHid, kHid, 1.0);
  fill(kOffW3, kOut, kHid, 0.1);  // the network starts near the guided interpolation
  std::vector<float> z(kParams, 0.0f);
  VCT_HIP(hipMemcpyAsync(ctx->mlp_params.p, p.data(), kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->mlp_m1.p, z.data(), kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->mlp_m2.p, z.data(), kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->mlp_step = 0;
  ctx->mlp_ready = true;
  return mlp_pack(ctx);
}

int vct_mlp_set_params(vct_ctx ctx, const float* params, uint32_t n) {
  MLP_CTX(ctx);
  if (!params || n != (uint32_t)kParams) return vct_set_error(VCT_ERR_INVALID_ARG, "need VCT_MLP_PARAMS floats");
  int rc;
  if ((rc = mlp_ensure(ctx))) return rc;
  std::vector<float> z(kParams, 0.0f);
  VCT_HIP(hipMemcpyAsync(ctx->mlp_params.p, params, kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->mlp_m1.p, z.data(), kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->mlp_m2.p, z.data(), kParams * 4, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->mlp_step = 0;
  ctx->mlp_ready = true;
  return mlp_pack(ctx);
}

int vct_mlp_get_params(vct_ctx ctx, float* params, uint32_t n) {
  MLP_CTX(ctx);
  if (!params || n != (uint32_t)kParams) return vct_set_error(VCT_ERR_INVALID_ARG, "need VCT_MLP_PARAMS floats");
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(params, ctx->mlp_params.p, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

int vct_mlp_forward(vct_ctx ctx, const float* features, uint32_t n, float* out) {
  MLP_CTX(ctx);
  if (!features || !out || n == 0) return vct_set_error(VCT_ERR_INVALID_ARG, "bad arguments");
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  if ((rc = upload_features(ctx, features, n, ctx->train_X))) return rc;
  if ((rc = vct_ensure(ctx->train_Y, (size_t)n * 12, ctx))) return rc;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, mlp_dev(ctx),
                     (const uint16_t*)ctx->train_X.p, n, (float*)ctx->train_Y.p);
  VCT_HIP(hipGetLastError());
  VCT_HIP(hipMemcpyAsync(out, ctx->train_Y.p, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->train_n = 0;  // the training set buffers were reused
  return VCT_OK;
}

int vct_mlp_forward_noalloc(vct_ctx ctx, const float* features, uint32_t n, float* out, void* stream) {
  MLP_CTX(ctx);
  if (!features || !out || n == 0) return vct_set_error(VCT_ERR_INVALID_ARG, "bad arguments");
  if (n > (uint32_t)VCT_MLP_NOALLOC_MAX_ROWS)
    return vct_set_error(VCT_ERR_LIMIT, "vct_mlp_forward_noalloc takes at most VCT_MLP_NOALLOC_MAX_ROWS rows");
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  uint16_t* X = ctx->device < kScratchDevices ? g_scratch_ptr[ctx->device] : nullptr;
  if (!X) return vct_set_error(VCT_ERR_STATE, "staging array of the device not set up");
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipLaunchKernelGGL(mlp_stage_kernel, dim3((unsigned)(((uint64_t)n * 8u + 255u) / 256u)), dim3(256), 0, st,
                     features, n, X);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, st, mlp_dev(ctx),
                     (const uint16_t*)X, n, out);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

int vct_mlp_loss_grad(vct_ctx ctx, const float* features, const float* albedo, const float* target, uint32_t n,
                      float* grad, float* loss) {
  MLP_CTX(ctx);
  if (!features || !albedo || !target || n == 0) return vct_set_error(VCT_ERR_INVALID_ARG, "bad arguments");
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  if ((rc = upload_features(ctx, features, n, ctx->train_X))) return rc;
  if ((rc = vct_ensure(ctx->train_A, (size_t)n * 12, ctx))) return rc;
  if ((rc = vct_ensure(ctx->train_Y, (size_t)n * 12, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->train_A.p, albedo, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->train_Y.p, target, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  ctx->train_n = 0;
  return mlp_train_on(ctx, (const uint16_t*)ctx->train_X.p, (const float*)ctx->train_A.p,
                      (const float*)ctx->train_Y.p, n, 0.0f, loss, grad);
}

int vct_upsample_set_samples(vct_ctx ctx, const float* features, const float* albedo, const float* target,
                             uint32_t n) {
  MLP_CTX(ctx);
  if (!features || !albedo || !target || n == 0) return vct_set_error(VCT_ERR_INVALID_ARG, "bad arguments");
  int rc;
  if ((rc = upload_features(ctx, features, n, ctx->train_X))) return rc;
  if ((rc = vct_ensure(ctx->train_A, (size_t)n * 12, ctx))) return rc;
  if ((rc = vct_ensure(ctx->train_Y, (size_t)n * 12, ctx))) return rc;
  if ((rc = vct_ensure(ctx->train_valid, (size_t)n * 4, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->train_A.p, albedo, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipMemcpyAsync(ctx->train_Y.p, target, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->train_n = n;
  ctx->train_cap = std::min(ctx->train_X.cap / 128, std::min(ctx->train_A.cap / 12, ctx->train_Y.cap / 12));
  return VCT_OK;
}

int vct_upsample_clear_samples(vct_ctx ctx) {
  MLP_CTX(ctx);
  ctx->train_n = 0;
  return VCT_OK;
}

int vct_upsample_sample_count(vct_ctx ctx, uint64_t* n) {
  MLP_CTX(ctx);
  if (!n) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  *n = ctx->train_n;
  return VCT_OK;
}

int vct_upsample_collect(vct_ctx ctx, uint32_t nsamples, uint64_t seed) {
  MLP_CTX(ctx);
  if (nsamples == 0) return VCT_OK;
  const int W = ctx->width, H = ctx->height;
  const size_t npix = (size_t)W * H;
  if (!ctx->g_pos.p || ctx->g_pos.cap < npix * 16)
    return vct_set_error(VCT_ERR_STATE, "collect needs a G-buffer of the current viewport (vct_gbuffer first)");
  int rc;
  // full-resolution plain trace, demodulated: the target
  if ((rc = vct_ensure(ctx->full_target, npix * 16, ctx))) return rc;
  if ((rc = vct_trace_into(ctx, (const float4*)ctx->g_pos.p, (const float4*)ctx->g_nrm.p,
                           (float4*)ctx->full_target.p, W, H, 1, nullptr)))
    return rc;
  if ((rc = trace_lowres(ctx))) return rc;
  // random pixels (background ones get zero weight)
  std::vector<uint32_t> pix(nsamples);
  HostRng rng(seed);
  for (uint32_t i = 0; i < nsamples; ++i) pix[i] = (uint32_t)(rng.next() % npix);
  if ((rc = vct_ensure(ctx->sample_pix, (size_t)nsamples * 4, ctx))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->sample_pix.p, pix.data(), (size_t)nsamples * 4, hipMemcpyHostToDevice, ctx->stream));
  const size_t need = ctx->train_n + nsamples;
  if (need > ctx->train_cap) {
    size_t cap = std::max<size_t>(need, ctx->train_cap * 2);
    VctBuf nx, na, ny, nv;
    if ((rc = vct_ensure(nx, cap * 128, ctx))) return rc;
    if ((rc = vct_ensure(na, cap * 12, ctx))) return rc;
    if ((rc = vct_ensure(ny, cap * 12, ctx))) return rc;
    if ((rc = vct_ensure(nv, cap * 4, ctx))) return rc;
    if (ctx->train_n) {
      VCT_HIP(hipMemcpyAsync(nx.p, ctx->train_X.p, ctx->train_n * 128, hipMemcpyDeviceToDevice, ctx->stream));
      VCT_HIP(hipMemcpyAsync(na.p, ctx->train_A.p, ctx->train_n * 12, hipMemcpyDeviceToDevice, ctx->stream));
      VCT_HIP(hipMemcpyAsync(ny.p, ctx->train_Y.p, ctx->train_n * 12, hipMemcpyDeviceToDevice, ctx->stream));
    }
    VCT_HIP(hipStreamSynchronize(ctx->stream));
    VctBuf* olds[] = {&ctx->train_X, &ctx->train_A, &ctx->train_Y, &ctx->train_valid};
    for (VctBuf* b : olds)
      if (b->p) VCT_HIP(hipFree(b->p));
    ctx->train_X = nx;
    ctx->train_A = na;
    ctx->train_Y = ny;
    ctx->train_valid = nv;
    ctx->train_cap = cap;
  }
  FeatArgs F = feat_args(ctx);
  const size_t o = ctx->train_n;
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((nsamples + 255) / 256), dim3(256), 0, ctx->stream, F,
                     (const float4*)ctx->full_target.p, (const uint32_t*)ctx->sample_pix.p, nsamples,
                     (uint16_t*)ctx->train_X.p + o * 64, (float*)ctx->train_A.p + o * 3,
                     (float*)ctx->train_Y.p + o * 3, (uint32_t*)ctx->train_valid.p + o);
  VCT_HIP(hipGetLastError());
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  ctx->train_n = need;
  return VCT_OK;
}

int vct_upsample_train(vct_ctx ctx, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* last_loss) {
  MLP_CTX(ctx);
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  if (ctx->train_n < 64) return vct_set_error(VCT_ERR_STATE, "no training samples (vct_upsample_collect first)");
  if (batch < 64 || !(lr > 0.0f)) return vct_set_error(VCT_ERR_INVALID_ARG, "batch >= 64 and lr > 0 required");
  if (ctx->train_n > 0xffffffffull) return vct_set_error(VCT_ERR_LIMIT, "more than 2^32 - 1 training samples");
  const uint32_t n = (uint32_t)ctx->train_n;
  batch = std::min<uint32_t>(batch, n) & ~63u;
  // Minibatches walk one random order of the whole sample set (samples are
  // stored frame by frame, so a contiguous window would hold one camera pose).
  if ((rc = shuffle_rows(ctx, n, seed))) return rc;
  if ((rc = vct_ensure(ctx->batch_X, (size_t)batch * 128, ctx))) return rc;
  if ((rc = vct_ensure(ctx->batch_A, (size_t)batch * 12, ctx))) return rc;
  if ((rc = vct_ensure(ctx->batch_Y, (size_t)batch * 12, ctx))) return rc;
  for (uint32_t s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(mlp_batch_gather_kernel, dim3((batch * 8 + 255) / 256), dim3(256), 0, ctx->stream,
                       (const uint32_t*)ctx->train_perm.p, n, (uint64_t)s * batch, batch,
                       (const uint16_t*)ctx->train_X.p, (const float*)ctx->train_A.p, (const float*)ctx->train_Y.p,
                       (uint16_t*)ctx->batch_X.p, (float*)ctx->batch_A.p, (float*)ctx->batch_Y.p);
    VCT_HIP(hipGetLastError());
    const bool last = s + 1 == steps;
    if ((rc = mlp_train_on(ctx, (const uint16_t*)ctx->batch_X.p, (const float*)ctx->batch_A.p,
                           (const float*)ctx->batch_Y.p, batch, lr, last ? last_loss : nullptr, nullptr)))
      return rc;
  }
  return VCT_OK;
}

// The learned form of Renderer::render: voxelize, mip, full-resolution
// G-buffer, quarter-resolution cone trace, MLP upsampling to full resolution,
// per-tile check and full-resolution trace of the tiles that fail it.
int vct_render_learned(vct_ctx ctx) {
  MLP_CTX(ctx);
  int rc;
  if ((rc = need_mlp(ctx))) return rc;
  if ((rc = lowres_buffers(ctx))) return rc;
  ctx->front_lowres = true;  // the quarter-resolution G-buffer is gathered beside voxelize + mip
  rc = vct_frame_front(ctx);
  ctx->front_lowres = false;
  if (rc) return rc;
  const size_t npix = (size_t)ctx->width * ctx->height;
  if ((rc = vct_ensure(ctx->image, npix * 16, ctx))) return rc;
  const bool fallback = ctx->fb_psnr_db > 0.0f;
  if (fallback) {  // the probes beside the quarter-resolution trace and the MLP
    VCT_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    VCT_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    std::swap(ctx->stream, ctx->side);
    rc = refine_probes(ctx);
    std::swap(ctx->stream, ctx->side);
    if (rc) return rc;
    VCT_HIP(hipEventRecord(ctx->ev_join, ctx->side));
  }
  if ((rc = vct_pass_event(ctx, 3, false, ctx->stream)) || (rc = trace_lowres(ctx, true)) ||
      (rc = vct_pass_event(ctx, 3, true, ctx->stream)) || (rc = vct_pass_event(ctx, 4, false, ctx->stream)))
    return rc;
  // frame budget: the learned values at the probes first, the tiles to trace
  // in full picked from them, then the network only outside those tiles
  const bool budget = fallback && !ctx->fb_half;
  if (budget) {
    if ((rc = infer_probes(ctx))) return rc;
    VCT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    if ((rc = budget_select(ctx, false))) return rc;
    if ((rc = infer_full(ctx, (const uint32_t*)ctx->fb_flag.p))) return rc;
  } else if ((rc = infer_full(ctx))) {
    return rc;
  }
  if ((rc = vct_pass_event(ctx, 4, true, ctx->stream)) || (rc = vct_pass_event(ctx, 5, false, ctx->stream)))
    return rc;
  ctx->fb_ran = false;
  if (budget) {
    const uint32_t np = (uint32_t)(((ctx->width + kFbTile - 1) / kFbTile) * ((ctx->height + kFbTile - 1) / kFbTile)) *
                        kProbes;
    hipLaunchKernelGGL(probe_write_kernel, dim3((np + 255) / 256), dim3(256), 0, ctx->stream,
                       (const float4*)ctx->probe_val.p, (const uint32_t*)ctx->probe_pix.p, np,
                       (float4*)ctx->image.p);
    VCT_HIP(hipGetLastError());
    if ((rc = refine(ctx, true))) return rc;
    ctx->fb_ran = true;
  } else if (fallback) {
    VCT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    if ((rc = refine(ctx, false))) return rc;
    ctx->fb_ran = true;
  } else if ((rc = trace_listed_pixels(ctx))) {
    return rc;
  }
  if ((rc = vct_pass_event(ctx, 5, true, ctx->stream))) return rc;
  return vct_clear_queues(ctx);
}

int vct_upsample_set_fallback(vct_ctx ctx, float min_tile_psnr_db) {
  MLP_CTX(ctx);
  if (!(min_tile_psnr_db == min_tile_psnr_db)) return vct_set_error(VCT_ERR_INVALID_ARG, "NaN threshold");
  ctx->fb_psnr_db = min_tile_psnr_db;
  return VCT_OK;
}

int vct_upsample_traced_pixels(vct_ctx ctx, uint32_t* pixels) {
  MLP_CTX(ctx);
  if (!pixels) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  *pixels = 0;
  if (ctx->fb_pix.p) {
    VCT_HIP(hipMemcpyAsync(pixels, ctx->fb_pix.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    VCT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return VCT_OK;
}

int vct_upsample_set_refine(vct_ctx ctx, int half_resolution) {
  MLP_CTX(ctx);
  ctx->fb_half = half_resolution != 0;
  return VCT_OK;
}

int vct_upsample_stats(vct_ctx ctx, uint32_t out[6]) {
  MLP_CTX(ctx);
  if (!out) return vct_set_error(VCT_ERR_INVALID_ARG, "null out");
  for (int i = 0; i < 6; ++i) out[i] = 0;
  out[0] = (uint32_t)(((ctx->width + kFbTile - 1) / kFbTile) * ((ctx->height + kFbTile - 1) / kFbTile));
  auto rd = [&](const VctBuf& b, uint32_t& v) -> int {
    if (!b.p) return VCT_OK;
    VCT_HIP(hipMemcpyAsync(&v, b.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    return VCT_OK;
  };
  int rc;
  if ((rc = rd(ctx->fb_pix, out[3]))) return rc;
  if (ctx->fb_ran) {
    if ((rc = rd(ctx->fb_list, out[1]))) return rc;
    if (ctx->fb_half_ran) {
      if ((rc = rd(ctx->fb_list2, out[2]))) return rc;
      if ((rc = rd(ctx->fb_pix2, out[4]))) return rc;
    }
  }
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->fb_ran && !ctx->fb_half_ran) out[2] = out[1];  // every failing tile was traced in full
  if (ctx->fb_ran) out[5] = out[0] * kProbes + (ctx->fb_half_ran ? out[1] * kProbes : 0u);
  return VCT_OK;
}

int vct_upsample_fallback_stats(vct_ctx ctx, uint32_t* tiles, uint32_t* fallback_tiles) {
  uint32_t st[6];
  int rc = vct_upsample_stats(ctx, st);
  if (rc) return rc;
  if (tiles) *tiles = st[0];
  if (fallback_tiles) *fallback_tiles = st[2];
  return VCT_OK;
}

// Quarter-resolution buffer of the last learned frame (tests, tools).
int vct_upsample_lowres_download(vct_ctx ctx, float* rgba, size_t pixels) {
  MLP_CTX(ctx);
  if (!rgba || pixels != (size_t)(ctx->width / 4) * (ctx->height / 4))
    return vct_set_error(VCT_ERR_INVALID_ARG, "pixel count mismatch");
  if (!ctx->low_img.p) return vct_set_error(VCT_ERR_STATE, "no low-resolution image yet");
  VCT_HIP(hipMemcpyAsync(rgba, ctx->low_img.p, pixels * 16, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  return VCT_OK;
}

}  // extern "C"
