// vct_gbuffer.hip - per-pixel surface buffer for the cone marcher.
//
// The reference gets the surface under each pixel by forward rendering with
// depth test: voxel_cone_tracing.vert:22-28 (world = model * p, clip =
// projection * view * world, normal by the inverse-transpose model matrix)
// rasterised at W x H with GL_LESS depth (renderer.cpp:355-390).  Here:
//  1. count  (lane per triangle): clip, near-plane reject, 1/256-pixel snap,
//     pixel bounding box in 2x2 quads.
//  2. exclusive scan of the candidate counts.
//  3. raster (lane per candidate pixel): exact 64-bit edge functions with the
//     top-left rule, screen-linear depth, 64-bit atomicMin of
//     (depth bits << 32 | triangle id).  Equal depths resolve to the lower
//     triangle id, which is the draw-order winner of GL_LESS.
//  4. resolve (lane per pixel): perspective-correct world position and
//     normal of the winning triangle, linear view depth, material id.
// Triangles with a vertex behind the near plane are skipped (not clipped).
#include <cmath>
#include <cstring>

#include "vct_internal.h"

namespace vct {

struct GDraw {
  const float* pos;
  const float* nrm;
  const uint32_t* mat;
  const uint32_t* perm;  // stored -> original triangle index
  const uint32_t* inv;   // original -> stored
  uint32_t tri_base;
  uint32_t ntri;
  float M[16];
  float N[9];
};

constexpr int kMaxGDraws = 16;

// Passed by value: no host staging buffer outlives the launch.
struct GParams {
  GDraw draws[kMaxGDraws];
  float PV[16];  // projection * view
  int ndraw;
  uint32_t ntri;
  int W, H;
};

struct GSetup {
  float clip[3][4];
  int64_t X[3], Y[3];
  int64_t area;
  bool swap, ok;
  int qx0, qy0, qw, qh;
};

__device__ __forceinline__ int g_find_draw(const GParams& p, uint32_t t) {
  int d = 0;
  for (int i = 1; i < p.ndraw; ++i)
    if (t >= p.draws[i].tri_base) d = i;
  return d;
}

__device__ __forceinline__ void world_of(const GDraw& dr, const float* p, float* w) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; ++k) {
    float t = dr.M[0 + k] * p[0];
    t = t + dr.M[4 + k] * p[1];
    t = t + dr.M[8 + k] * p[2];
    t = t + dr.M[12 + k];
    w[k] = t;
  }
}

__device__ __forceinline__ void g_setup(const GDraw& dr, const float* PV, uint32_t lt, int W, int H, GSetup& s) {
#pragma clang fp contract(off)
  const float* p = dr.pos + (size_t)lt * 9;
  bool ok = true;
  for (int v = 0; v < 3; ++v) {
    float w[3];
    world_of(dr, p + 3 * v, w);
    for (int k = 0; k < 4; ++k) {
      float t = PV[0 + k] * w[0];
      t = t + PV[4 + k] * w[1];
      t = t + PV[8 + k] * w[2];
      t = t + PV[12 + k];
      s.clip[v][k] = t;
    }
    const float cw = s.clip[v][3];
    if (!(cw > 0.0f) || s.clip[v][2] < -cw) ok = false;
  }
  const float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    float xn = s.clip[v][0] / s.clip[v][3], yn = s.clip[v][1] / s.clip[v][3];
    if (!(fabsf(xn) < 256.0f && fabsf(yn) < 256.0f)) {
      ok = false;
      xn = 0.0f;
      yn = 0.0f;
    }
    const float xw = xn * hw + hw, yw = yn * hh + hh;
    s.X[v] = (int64_t)__builtin_rintf((xw - 0.5f) * 256.0f);
    s.Y[v] = (int64_t)__builtin_rintf((yw - 0.5f) * 256.0f);
  }
  s.ok = false;
  if (!ok) return;
  int64_t area = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
  if (area == 0) return;
  s.swap = area < 0;
  s.area = area < 0 ? -area : area;
  int64_t xmin = min(s.X[0], min(s.X[1], s.X[2])), xmax = max(s.X[0], max(s.X[1], s.X[2]));
  int64_t ymin = min(s.Y[0], min(s.Y[1], s.Y[2])), ymax = max(s.Y[0], max(s.Y[1], s.Y[2]));
  int64_t imin = (xmin + 255) >> 8, imax = xmax >> 8, jmin = (ymin + 255) >> 8, jmax = ymax >> 8;
  imin = imin < 0 ? 0 : imin;
  jmin = jmin < 0 ? 0 : jmin;
  imax = imax > W - 1 ? W - 1 : imax;
  jmax = jmax > H - 1 ? H - 1 : jmax;
  if (imin > imax || jmin > jmax) return;
  s.qx0 = (int)(imin & ~1ll);
  s.qy0 = (int)(jmin & ~1ll);
  s.qw = (int)((imax - s.qx0) >> 1) + 1;
  s.qh = (int)((jmax - s.qy0) >> 1) + 1;
  s.ok = true;
}

__device__ __forceinline__ bool g_top_left(int64_t dx, int64_t dy) { return dy < 0 || (dy == 0 && dx < 0); }

// Barycentric weights (of vertices 0,1,2) of pixel (i,j); false if outside.
__device__ __forceinline__ bool g_cover(const GSetup& s, int i, int j, float& b0, float& b1, float& b2) {
#pragma clang fp contract(off)
  const int64_t px = (int64_t)i * 256, py = (int64_t)j * 256;
  const int ia = 0, ib = s.swap ? 2 : 1, ic = s.swap ? 1 : 2;
  const int64_t ax = s.X[ia], ay = s.Y[ia], bx = s.X[ib], by = s.Y[ib], cx = s.X[ic], cy = s.Y[ic];
  int64_t wa = (cx - bx) * (py - by) - (cy - by) * (px - bx);
  int64_t wb = (ax - cx) * (py - cy) - (ay - cy) * (px - cx);
  int64_t wc = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
  if (wa < 0 || wb < 0 || wc < 0) return false;
  if (wa == 0 && !g_top_left(cx - bx, cy - by)) return false;
  if (wb == 0 && !g_top_left(ax - cx, ay - cy)) return false;
  if (wc == 0 && !g_top_left(bx - ax, by - ay)) return false;
  int64_t w[3];
  w[ia] = wa;
  w[ib] = wb;
  w[ic] = wc;
  const float fa = (float)s.area;
  b0 = (float)w[0] / fa;
  b1 = (float)w[1] / fa;
  b2 = (float)w[2] / fa;
  return true;
}

// Per-triangle record of the count pass, read by the triangle's candidate lanes.
struct alignas(16) GRec {
  int64_t area;
  int32_t X[3], Y[3];
  float z[3];  // window depth of each vertex
  int32_t qx0, qy0, qw;
  uint32_t swap;
  uint32_t otri;  // original triangle index in the frame (GL draw order)
};
static_assert(sizeof(GRec) == 64, "GRec must be 64 bytes");

__global__ __launch_bounds__(256) void g_count_kernel(GParams p, uint32_t* counts, GRec* __restrict__ recs) {
#pragma clang fp contract(off)
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.ntri) return;
  const int d = g_find_draw(p, t);
  GSetup s;
  g_setup(p.draws[d], p.PV, t - p.draws[d].tri_base, p.W, p.H, s);
  counts[t] = s.ok ? (uint32_t)(4 * s.qw * s.qh) : 0u;
  if (!s.ok) return;
  GRec r;
  r.area = s.area;
  for (int v = 0; v < 3; ++v) {
    r.X[v] = (int32_t)s.X[v];
    r.Y[v] = (int32_t)s.Y[v];
    r.z[v] = (s.clip[v][2] / s.clip[v][3]) * 0.5f + 0.5f;
  }
  r.qx0 = s.qx0;
  r.qy0 = s.qy0;
  r.qw = s.qw;
  r.swap = s.swap ? 1u : 0u;
  r.otri = p.draws[d].tri_base + p.draws[d].perm[t - p.draws[d].tri_base];
  recs[t] = r;
}

__global__ __launch_bounds__(256) void g_raster_kernel(uint32_t ntri, int W, int H, const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ d_total,
                                                       const GRec* __restrict__ recs, unsigned long long* vis) {
#pragma clang fp contract(off)
  __shared__ uint32_t s_range[2];
  const uint32_t total = *d_total;
  const uint32_t n = ntri;
  for (uint32_t base = blockIdx.x * 256u; base < total; base += gridDim.x * 256u) {
    const uint32_t last = min(base + 255u, total - 1u);
    if (threadIdx.x == 0 || threadIdx.x == 255) {
      const uint32_t key = threadIdx.x == 0 ? base : last;
      uint32_t lo = 0, hi = n;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= key) lo = mid;
        else hi = mid;
      }
      s_range[threadIdx.x == 0 ? 0 : 1] = lo;
    }
    __syncthreads();
    const uint32_t c = base + threadIdx.x;
    if (c < total) {
      uint32_t lo = s_range[0], hi = s_range[1] + 1;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= c) lo = mid;
        else hi = mid;
      }
      const uint32_t t = lo;
      const GRec r = recs[t];
      GSetup s;
      s.area = r.area;
      s.swap = r.swap != 0u;
      for (int v = 0; v < 3; ++v) {
        s.X[v] = r.X[v];
        s.Y[v] = r.Y[v];
      }
      const uint32_t lc = c - off[t];
      const uint32_t quad = lc >> 2, sub = lc & 3u;
      const int i = r.qx0 + 2 * (int)(quad % (uint32_t)r.qw) + (int)(sub & 1u);
      const int j = r.qy0 + 2 * (int)(quad / (uint32_t)r.qw) + (int)(sub >> 1);
      float b0, b1, b2;
      if (i < W && j < H && g_cover(s, i, j, b0, b1, b2)) {
        const float depth = b0 * r.z[0] + b1 * r.z[1] + b2 * r.z[2];
        if (depth >= 0.0f && depth <= 1.0f) {
          const unsigned long long key =
              ((unsigned long long)__builtin_bit_cast(uint32_t, depth) << 32) | (unsigned long long)r.otri;
          unsigned long long* dst = &vis[(size_t)j * W + i];
          // early depth test: skip the atomic when a nearer surface is already in
          // (a stale read only costs a redundant atomic, never a wrong result)
          if (key < __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(dst, key);
        }
      }
    }
    __syncthreads();
  }
}

// Reads the winner of each pixel and resets the entry to empty for the next
// frame (so the visibility buffer never needs a separate clear).
__global__ __launch_bounds__(256) void g_resolve_kernel(GParams p, unsigned long long* vis,
                                                        float4* __restrict__ gpos, float4* __restrict__ gnrm) {
#pragma clang fp contract(off)
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint32_t)(p.W * p.H)) return;
  const unsigned long long key = vis[idx];
  vis[idx] = ~0ull;
  if (key == ~0ull) {
    gpos[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    gnrm[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const uint32_t t = (uint32_t)(key & 0xffffffffu);  // original index
  const int d = g_find_draw(p, t);
  const GDraw& dr = p.draws[d];
  const uint32_t lt = dr.inv[t - dr.tri_base];       // stored index
  GSetup s;
  g_setup(dr, p.PV, lt, p.W, p.H, s);
  const int i = (int)(idx % (uint32_t)p.W), j = (int)(idx / (uint32_t)p.W);
  float b[3];
  if (!g_cover(s, i, j, b[0], b[1], b[2])) {  // cannot happen: same arithmetic as the raster
    b[0] = b[1] = b[2] = 1.0f / 3.0f;
  }
  // perspective-correct weights: (b_i / w_i) / sum(b_k / w_k)
  float q[3], qs = 0.0f;
  for (int v = 0; v < 3; ++v) {
    q[v] = b[v] / s.clip[v][3];
    qs = qs + q[v];
  }
  for (int v = 0; v < 3; ++v) q[v] = q[v] / qs;
  const float* pp = dr.pos + (size_t)lt * 9;
  const float* nn = dr.nrm + (size_t)lt * 9;
  float wp[3] = {0.f, 0.f, 0.f}, n[3] = {0.f, 0.f, 0.f};
  for (int v = 0; v < 3; ++v) {
    float w[3];
    world_of(dr, pp + 3 * v, w);
    float t3[3];
    for (int k = 0; k < 3; ++k) {
      float a = dr.N[0 + k] * nn[v * 3 + 0];
      a = a + dr.N[3 + k] * nn[v * 3 + 1];
      a = a + dr.N[6 + k] * nn[v * 3 + 2];
      t3[k] = a;
    }
    const float l = sqrtf(t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2]);
    for (int k = 0; k < 3; ++k) {
      wp[k] = wp[k] + q[v] * w[k];
      n[k] = n[k] + q[v] * (t3[k] / l);
    }
  }
  const float lin_depth = 1.0f / qs;  // interpolated clip w
  gpos[idx] = make_float4(wp[0], wp[1], wp[2], lin_depth);
  gnrm[idx] = make_float4(n[0], n[1], n[2], __builtin_bit_cast(float, dr.mat[lt]));
}

__global__ void fill_u64_kernel(unsigned long long* p, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace vct

using namespace vct;

static void mat4_mul(const float* A, const float* B, float* C) {
  // C = A * B, column-major, each element summed k = 0..3 in float
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      float s = A[0 * 4 + r] * B[c * 4 + 0];
      s = s + A[1 * 4 + r] * B[c * 4 + 1];
      s = s + A[2 * 4 + r] * B[c * 4 + 2];
      s = s + A[3 * 4 + r] * B[c * 4 + 3];
      C[c * 4 + r] = s;
    }
}

static void normal_matrix3(const float* M, float* N) {
  double a[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[r][c] = M[c * 4 + r];
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) N[c * 3 + r] = (float)inv[c][r];
}

int vct_launch_gbuffer(vct_context* ctx) {
  const int W = ctx->width, H = ctx->height;
  const size_t npix = (size_t)W * H;
  int rc;
  if ((rc = vct_ensure(ctx->g_pos, npix * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->g_nrm, npix * 16, ctx))) return rc;
  const bool fresh_vis = !ctx->vis.p || ctx->vis.cap < npix * 8;
  if ((rc = vct_ensure(ctx->vis, npix * 8, ctx))) return rc;
  if (fresh_vis || ctx->vis_pixels != npix) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1024), dim3(256), 0, ctx->stream, (unsigned long long*)ctx->vis.p,
                       (uint64_t)npix, ~0ull);
    ctx->vis_pixels = npix;
  }
  if ((rc = vct_ensure(ctx->counters, 256, ctx))) return rc;
  uint32_t* counters = (uint32_t*)ctx->counters.p;

  GParams p;
  memset(&p, 0, sizeof(p));
  uint32_t ntri = 0;
  mat4_mul(ctx->proj, ctx->view, p.PV);
  for (uint32_t id : ctx->queue) {
    if (id >= ctx->models.size() || !ctx->models[id].ntri) continue;
    if (p.ndraw == kMaxGDraws) return vct_set_error(VCT_ERR_LIMIT, "at most 16 queued models per frame");
    const VctModel& m = ctx->models[id];
    GDraw& d = p.draws[p.ndraw++];
    d.pos = m.d_pos;
    d.nrm = m.d_nrm;
    d.mat = m.d_mat;
    d.perm = m.d_perm;
    d.inv = m.d_inv;
    d.tri_base = ntri;
    d.ntri = m.ntri;
    memcpy(d.M, m.xform, 64);
    normal_matrix3(m.xform, d.N);
    ntri += m.ntri;
  }
  p.ntri = ntri;
  p.W = W;
  p.H = H;
  if (ntri) {
    if ((rc = vct_ensure(ctx->g_count, (size_t)ntri * 4 + 4, ctx))) return rc;
    if ((rc = vct_ensure(ctx->g_off, (size_t)ntri * 4 + 4, ctx))) return rc;
    if ((rc = vct_ensure(ctx->g_recs, (size_t)ntri * sizeof(GRec), ctx))) return rc;
    hipLaunchKernelGGL(g_count_kernel, dim3((ntri + 255) / 256), dim3(256), 0, ctx->stream, p,
                       (uint32_t*)ctx->g_count.p, (GRec*)ctx->g_recs.p);
    VCT_HIP(hipGetLastError());
    if ((rc = vct_launch_scan_u32(ctx, (uint32_t*)ctx->g_count.p, (uint32_t*)ctx->g_off.p, ntri, counters + 4)))
      return rc;
    hipLaunchKernelGGL(g_raster_kernel, dim3(256 * 8), dim3(256), 0, ctx->stream, ntri, W, H,
                       (const uint32_t*)ctx->g_off.p, (const uint32_t*)(counters + 4), (const GRec*)ctx->g_recs.p,
                       (unsigned long long*)ctx->vis.p);
    VCT_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(g_resolve_kernel, dim3((uint32_t)((npix + 255) / 256)), dim3(256), 0, ctx->stream, p,
                     (unsigned long long*)ctx->vis.p, (float4*)ctx->g_pos.p, (float4*)ctx->g_nrm.p);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}
