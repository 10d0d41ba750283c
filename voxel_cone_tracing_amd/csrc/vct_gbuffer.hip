// vct_gbuffer.hip - per-pixel surface buffer for the cone marcher.
//
// The reference gets the surface under each pixel by forward rendering with
// depth test: voxel_cone_tracing.vert:22-28 (world = model * p, clip =
// projection * view * world, normal by the inverse-transpose model matrix)
// rasterised at W x H with GL_LESS depth (renderer.cpp:355-390).  Here, a
// tiled rasteriser with the depth buffer in LDS:
//  1. setup + bin count (lane per triangle): clip coordinates, GL's primitive
//     clipping (below), 1/256-pixel snap, pixel bounding box; count into the
//     32x32 tiles the box overlaps (one atomic per distinct tile per wave).
//  2. exclusive scan of the tile counts.
//  3. bin write: (triangle, fan triangle) entries into per-tile lists.
//  4. tile pass (workgroup per tile): exact 64-bit edge functions with the
//     top-left rule, screen-linear depth, LDS atomicMin of
//     (depth bits << 32 | triangle << 3 | fan triangle); then
//     perspective-correct world position and normal of each pixel's winner,
//     linear view depth and material id, written once per pixel.
// Clipping: GL clips every primitive against the view volume.  A triangle
// whose vertices all lie in front of the near plane (w > 0, z >= -w) and
// inside a guard band of +-256 NDC units is drawn as it is; any other is
// clipped in clip space against the near plane and the guard-band planes
// x, y = +-255 w (Sutherland-Hodgman, new vertices interpolated from the
// inside end of their edge) and the polygon left is drawn as the fan
// (0, k + 1, k + 2).  Its vertices carry weights over the original three, so
// attributes resolve against the original vertices.  Clipped triangles are
// rare (they pass near the eye), so their polygon is recomputed where it is
// needed instead of stored.
#include <cmath>
#include <cstring>

#include <hipcub/hipcub.hpp>

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

// Passed by value; the draw table lives in device memory (any number of
// queued models).
struct GParams {
  const GDraw* draws;
  int ndraw;
  float PV[16];  // projection * view
  uint32_t ntri;
  int W, H;
};

struct GSetup {
  float clip[3][4];
  int64_t X[3], Y[3];
  int64_t area;
  bool swap, ok;
  int i0, j0, i1, j1;  // pixel bounding box, inclusive
};

// The draw of frame triangle t (stored or original numbering: both count
// from the draw's tri_base): the last draw whose tri_base is <= t.
__device__ __forceinline__ int g_find_draw(const GParams& p, uint32_t t) {
  int lo = 0, hi = p.ndraw - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.draws[mid].tri_base <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
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

// Clip coordinates of the three vertices of stored triangle lt of a draw.
__device__ __forceinline__ void g_clip_coords(const GDraw& dr, const float* PV, uint32_t lt, float (&clip)[3][4]) {
#pragma clang fp contract(off)
  const float* p = dr.pos + (size_t)lt * 9;
  for (int v = 0; v < 3; ++v) {
    float w[3];
    world_of(dr, p + 3 * v, w);
    for (int k = 0; k < 4; ++k) {
      float t = PV[0 + k] * w[0];
      t = t + PV[4 + k] * w[1];
      t = t + PV[8 + k] * w[2];
      t = t + PV[12 + k];
      clip[v][k] = t;
    }
  }
}

// True if the triangle is drawn unclipped: every vertex in front of the near
// plane and inside the +-256 NDC guard band.
__device__ __forceinline__ bool g_unclipped(const float (&clip)[3][4]) {
#pragma clang fp contract(off)
  bool in = true;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const float cw = clip[v][3];
    if (!(cw > 0.0f) || clip[v][2] < -cw) in = false;
    else if (!(fabsf(clip[v][0] / cw) < 256.0f && fabsf(clip[v][1] / cw) < 256.0f)) in = false;
  }
  return in;
}

// Snap, orientation and pixel bounding box of a triangle given by the clip
// coordinates of its vertices (every w > 0, NDC within the guard band).
__device__ __forceinline__ void g_setup_clip(GSetup& s, int W, int H) {
#pragma clang fp contract(off)
  const float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
  bool ok = true;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    float xn = s.clip[v][0] / s.clip[v][3], yn = s.clip[v][1] / s.clip[v][3];
    if (!(s.clip[v][3] > 0.0f) || !(fabsf(xn) < 256.0f && fabsf(yn) < 256.0f)) {
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
  s.i0 = (int)imin;
  s.j0 = (int)jmin;
  s.i1 = (int)imax;
  s.j1 = (int)jmax;
  s.ok = true;
}

// A vertex of a clipped triangle: clip coordinates and its weights over the
// triangle's three original vertices.
struct ClipVert {
  float c[4];
  float b[3];
};

__device__ __forceinline__ float g_plane_dist(int plane, const float* c) {
#pragma clang fp contract(off)
  const float g = 255.0f;
  switch (plane) {
    case 0: return c[2] + c[3];  // near: z >= -w
    case 1: return g * c[3] - c[0];
    case 2: return g * c[3] + c[0];
    case 3: return g * c[3] - c[1];
    default: return g * c[3] + c[1];
  }
}

// Clips the triangle (clip coordinates) against the near plane and the
// guard-band planes; returns the polygon's vertex count (0 or 3..8).
__device__ __noinline__ int g_clip_polygon(const float (&clip)[3][4], ClipVert (&P)[8]) {
#pragma clang fp contract(off)
  int n = 3;
  for (int v = 0; v < 3; ++v) {
    for (int k = 0; k < 4; ++k) P[v].c[k] = clip[v][k];
    for (int k = 0; k < 3; ++k) P[v].b[k] = k == v ? 1.0f : 0.0f;
  }
  for (int plane = 0; plane < 5 && n > 0; ++plane) {
    ClipVert out[8];
    int m = 0;
    for (int i = 0; i < n; ++i) {
      const ClipVert a = P[i];
      const ClipVert b = P[i + 1 < n ? i + 1 : 0];
      const float da = g_plane_dist(plane, a.c), db = g_plane_dist(plane, b.c);
      const bool ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina && m < 8) out[m++] = a;
      if (ina != inb && m < 8) {
        const ClipVert& in = ina ? a : b;
        const ClipVert& ot = ina ? b : a;
        const float din = ina ? da : db, dout = ina ? db : da;
        const float t = din / (din - dout);
        ClipVert r;
        for (int k = 0; k < 4; ++k) r.c[k] = in.c[k] + t * (ot.c[k] - in.c[k]);
        for (int k = 0; k < 3; ++k) r.b[k] = in.b[k] + t * (ot.b[k] - in.b[k]);
        out[m++] = r;
      }
    }
    n = m;
    for (int i = 0; i < m; ++i) P[i] = out[i];
  }
  return n < 3 ? 0 : n;
}

// Setup of fan triangle k of the clipped triangle; wts receives the weights
// of its three vertices over the original ones.
__device__ __noinline__ void g_setup_fan(const float (&clip)[3][4], int k, int W, int H, GSetup& s,
                                         float (&wts)[3][3]) {
  ClipVert P[8];
  const int n = g_clip_polygon(clip, P);
  s.ok = false;
  if (k + 2 >= n) return;
  const int idx[3] = {0, k + 1, k + 2};
  for (int v = 0; v < 3; ++v) {
    for (int c = 0; c < 4; ++c) s.clip[v][c] = P[idx[v]].c[c];
    for (int c = 0; c < 3; ++c) wts[v][c] = P[idx[v]].b[c];
  }
  g_setup_clip(s, W, H);
}

// Number of fan triangles of a clipped triangle.
__device__ __noinline__ int g_fan_count(const float (&clip)[3][4]) {
  ClipVert P[8];
  const int n = g_clip_polygon(clip, P);
  return n ? n - 2 : 0;
}

__device__ __forceinline__ bool g_top_left(int64_t dx, int64_t dy) { return dy < 0 || (dy == 0 && dx < 0); }

// Edge functions of a set-up triangle in CCW order (a, b, c) = (0, 1, 2) or
// (0, 2, 1): w_a(i, j) = wa0 + i * dia + j * dja (exact 64-bit integers at
// pixel centres, coordinates in 1/256 pixel), likewise for b and c.
struct GEdges {
  int64_t w[3];    // at the start pixel
  int64_t di[3];   // per column
  int64_t dj[3];   // per row
  int64_t ex[3], ey[3];  // edge vectors (for the top-left rule)
};

__device__ __forceinline__ void g_edges(const GSetup& s, int i, int j, GEdges& e) {
  const int64_t px = (int64_t)i * 256, py = (int64_t)j * 256;
  const int ia = 0, ib = s.swap ? 2 : 1, ic = s.swap ? 1 : 2;
  const int64_t ax = s.X[ia], ay = s.Y[ia], bx = s.X[ib], by = s.Y[ib], cx = s.X[ic], cy = s.Y[ic];
  e.w[0] = (cx - bx) * (py - by) - (cy - by) * (px - bx);
  e.w[1] = (ax - cx) * (py - cy) - (ay - cy) * (px - cx);
  e.w[2] = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
  e.di[0] = -(cy - by) * 256;
  e.di[1] = -(ay - cy) * 256;
  e.di[2] = -(by - ay) * 256;
  e.dj[0] = (cx - bx) * 256;
  e.dj[1] = (ax - cx) * 256;
  e.dj[2] = (bx - ax) * 256;
  e.ex[0] = cx - bx;
  e.ey[0] = cy - by;
  e.ex[1] = ax - cx;
  e.ey[1] = ay - cy;
  e.ex[2] = bx - ax;
  e.ey[2] = by - ay;
}

// Coverage (top-left rule) and barycentric weights of vertices 0, 1, 2 from
// the edge values (wa, wb, wc) of a pixel.
__device__ __forceinline__ bool g_cover_w(bool swap, float fa, const GEdges& e, int64_t wa, int64_t wb, int64_t wc,
                                          float& b0, float& b1, float& b2) {
#pragma clang fp contract(off)
  if (wa < 0 || wb < 0 || wc < 0) return false;
  if (wa == 0 && !g_top_left(e.ex[0], e.ey[0])) return false;
  if (wb == 0 && !g_top_left(e.ex[1], e.ey[1])) return false;
  if (wc == 0 && !g_top_left(e.ex[2], e.ey[2])) return false;
  const int64_t w0 = wa, w1 = swap ? wc : wb, w2 = swap ? wb : wc;
  b0 = (float)w0 / fa;
  b1 = (float)w1 / fa;
  b2 = (float)w2 / fa;
  return true;
}

// Barycentric weights (of vertices 0,1,2) of pixel (i,j); false if outside.
__device__ __forceinline__ bool g_cover(const GSetup& s, int i, int j, float& b0, float& b1, float& b2) {
  GEdges e;
  g_edges(s, i, j, e);
  return g_cover_w(s.swap, (float)s.area, e, e.w[0], e.w[1], e.w[2], b0, b1, b2);
}

// Per-triangle record of the bin pass, read by every tile the triangle overlaps.
struct alignas(16) GRec {
  int64_t area;
  int32_t X[3], Y[3];
  float z[3];                // window depth of each vertex
  int16_t i0, j0, i1, j1;    // pixel bounding box, inclusive; i0 > i1 when culled
  uint32_t swap;             // bit 0: vertices 1 and 2 swapped for CCW; bit 1: clipped (drawn as its fan records)
  uint32_t otri;             // original triangle index in the frame (GL draw order) << 3 | fan triangle
  uint32_t nfan;             // clipped: fan triangles, stored from fan record `area`
};
static_assert(sizeof(GRec) == 64, "GRec must be 64 bytes");
constexpr uint32_t kClipped = 2u;

// What the resolve of a fan triangle needs besides its record: the clip w of
// its vertices and their weights over the original three.
struct alignas(16) GFan {
  float w[3];
  float wts[3][3];
  float pad[4];
};
static_assert(sizeof(GFan) == 64, "GFan must be 64 bytes");

constexpr int kTile = 32;  // screen tile edge in pixels (one workgroup per tile)

__device__ __forceinline__ GSetup g_setup_of(const GRec& r) {
  GSetup s;
  s.area = r.area;
  s.swap = (r.swap & 1u) != 0u;
  for (int v = 0; v < 3; ++v) {
    s.X[v] = r.X[v];
    s.Y[v] = r.Y[v];
  }
  return s;
}

struct TileRange {
  int tx0, ty0, tw, n;
};

__device__ __forceinline__ TileRange tile_range(const GRec& r) {
  TileRange t;
  t.n = 0;
  t.tx0 = t.ty0 = t.tw = 0;
  if (r.i0 > r.i1) return t;
  t.tx0 = r.i0 / kTile;
  t.ty0 = r.j0 / kTile;
  t.tw = r.i1 / kTile - t.tx0 + 1;
  t.n = t.tw * (r.j1 / kTile - t.ty0 + 1);
  return t;
}

__device__ __forceinline__ void g_rec_from(const GSetup& s, GRec& r) {
#pragma clang fp contract(off)
  r.area = s.area;
  for (int v = 0; v < 3; ++v) {
    r.X[v] = (int32_t)s.X[v];
    r.Y[v] = (int32_t)s.Y[v];
    r.z[v] = s.ok ? (s.clip[v][2] / s.clip[v][3]) * 0.5f + 0.5f : 0.0f;
  }
  r.i0 = 1;
  r.i1 = 0;
  r.j0 = 1;
  r.j1 = 0;
  if (s.ok) {
    r.i0 = (int16_t)s.i0;
    r.j0 = (int16_t)s.j0;
    r.i1 = (int16_t)s.i1;
    r.j1 = (int16_t)s.j1;
  }
  r.swap = s.swap ? 1u : 0u;
}

// The record (and resolve data) of fan triangle k of clipped stored triangle t.
__device__ __noinline__ GRec g_fan_record(const GParams& p, uint32_t t, uint32_t k, GFan* fan) {
  const int d = g_find_draw(p, t);
  const GDraw& dr = p.draws[d];
  float clip[3][4];
  g_clip_coords(dr, p.PV, t - dr.tri_base, clip);
  GSetup s;
  float wts[3][3];
  g_setup_fan(clip, (int)k, p.W, p.H, s, wts);
  GRec r;
  g_rec_from(s, r);
  r.otri = ((dr.tri_base + dr.perm[t - dr.tri_base]) << 3) | k;
  r.nfan = 0;
  if (fan) {
    for (int v = 0; v < 3; ++v) {
      fan->w[v] = s.clip[v][3];
      for (int o = 0; o < 3; ++o) fan->wts[v][o] = wts[v][o];
    }
    for (int q = 0; q < 4; ++q) fan->pad[q] = 0.0f;
  }
  return r;
}

// 1. setup (lane per stored triangle): clip, snap, bounding box; count the
//    triangle into every 32x32 tile its bounding box overlaps.  A triangle that
//    needs clipping is only flagged and listed here: its fans are counted and
//    written by g_fan_count_kernel / g_fan_write_kernel, so that the polygon
//    code (private arrays, calls) stays out of this kernel.
__global__ __launch_bounds__(256) void g_bin_count_kernel(GParams p, GRec* __restrict__ recs, uint32_t* tile_count,
                                                          int tiles_x, uint32_t* clip_list, uint32_t* clip_count) {
#pragma clang fp contract(off)
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  GRec r;
  r.i0 = 1;
  r.i1 = 0;
  r.j0 = 1;
  r.j1 = 0;
  bool clipped = false;
  if (t < p.ntri) {
    const int d = g_find_draw(p, t);
    const GDraw& dr = p.draws[d];
    float clip[3][4];
    g_clip_coords(dr, p.PV, t - dr.tri_base, clip);
    GSetup s;
    if (g_unclipped(clip)) {
      for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 4; ++c) s.clip[v][c] = clip[v][c];
      g_setup_clip(s, p.W, p.H);
      g_rec_from(s, r);
    } else {
      clipped = true;
      r.area = 0;
      r.swap = kClipped;
    }
    r.nfan = 0;
    r.otri = (dr.tri_base + dr.perm[t - dr.tri_base]) << 3;
    recs[t] = r;
  }
  const uint32_t slot = wave_add_one(clip_count, 0u, clipped);
  if (clipped) clip_list[slot] = t;
  const TileRange tr = tile_range(r);
  for (int k = 0; __ballot(k < tr.n) != 0ull; ++k) {
    const bool act = k < tr.n;
    const uint32_t key = act ? (uint32_t)((tr.ty0 + k / tr.tw) * tiles_x + tr.tx0 + k % tr.tw) : 0u;
    wave_add_one(tile_count, key, act);
  }
}

// 1b. the listed clipped triangles (a fixed grid striding over the device-side
//     count): each takes its fan records' range and counts its fans' tiles.
__global__ __launch_bounds__(256) void g_fan_count_kernel(GParams p, GRec* __restrict__ recs,
                                                          const uint32_t* __restrict__ clip_list,
                                                          const uint32_t* __restrict__ clip_count,
                                                          uint32_t* tile_count, int tiles_x, uint32_t* fan_total) {
  const uint32_t n = *clip_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = clip_list[i];
    const int d = g_find_draw(p, t);
    float clip[3][4];
    g_clip_coords(p.draws[d], p.PV, t - p.draws[d].tri_base, clip);
    const int nfan = g_fan_count(clip);
    recs[t].area = nfan ? (int64_t)atomicAdd(fan_total, (uint32_t)nfan) : 0;  // first fan record
    recs[t].nfan = (uint32_t)nfan;
    for (int f = 0; f < nfan; ++f) {
      const TileRange tr = tile_range(g_fan_record(p, t, (uint32_t)f, nullptr));
      for (int k = 0; k < tr.n; ++k) atomicAdd(tile_count + (tr.ty0 + k / tr.tw) * tiles_x + tr.tx0 + k % tr.tw, 1u);
    }
  }
}

// 3. list write: the record index of each stored triangle into each overlapped
//    tile's list (clipped triangles have no tiles here).
__global__ __launch_bounds__(256) void g_bin_write_kernel(uint32_t ntri, const GRec* __restrict__ recs,
                                                          const uint32_t* __restrict__ tile_off, uint32_t* cursor,
                                                          int tiles_x, uint32_t* __restrict__ list) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  GRec r;
  r.i0 = 1;
  r.i1 = 0;
  if (t < ntri) r = recs[t];
  const TileRange tr = tile_range(r);
  for (int k = 0; __ballot(k < tr.n) != 0ull; ++k) {
    const bool act = k < tr.n;
    const uint32_t key = act ? (uint32_t)((tr.ty0 + k / tr.tw) * tiles_x + tr.tx0 + k % tr.tw) : 0u;
    const uint32_t slot = wave_add_one(cursor, key, act);
    if (act) list[tile_off[key] + slot] = t;
  }
}

// 3b. the fan records of the clipped triangles and their list entries
//     (ntri + fan record).
__global__ __launch_bounds__(256) void g_fan_write_kernel(GParams p, const GRec* __restrict__ recs,
                                                          const uint32_t* __restrict__ clip_list,
                                                          const uint32_t* __restrict__ clip_count,
                                                          const uint32_t* __restrict__ tile_off, uint32_t* cursor,
                                                          int tiles_x, uint32_t* __restrict__ list,
                                                          GRec* __restrict__ fan_recs, GFan* __restrict__ fans) {
  const uint32_t n = *clip_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = clip_list[i];
    const uint32_t base = (uint32_t)recs[t].area, nfan = recs[t].nfan;
    for (uint32_t f = 0; f < nfan; ++f) {
      GFan fd;
      const GRec fr = g_fan_record(p, t, f, &fd);
      fan_recs[base + f] = fr;
      fans[base + f] = fd;
      const TileRange tr = tile_range(fr);
      for (int k = 0; k < tr.n; ++k) {
        const uint32_t key = (uint32_t)((tr.ty0 + k / tr.tw) * tiles_x + tr.tx0 + k % tr.tw);
        list[tile_off[key] + atomicAdd(cursor + key, 1u)] = p.ntri + base + f;
      }
    }
  }
}

// Perspective-correct interpolation weights q of the original vertices of
// triangle id (frame triangle << 3 | fan triangle) at pixel (i, j), as the
// raster drew it: its draw, stored index, and sum(b_k / w_k) (1 / qs is the
// interpolated clip w).
__device__ __forceinline__ const GDraw& g_weights(const GParams& p, const GRec* __restrict__ recs,
                                                  const GRec* __restrict__ fan_recs, const GFan* __restrict__ fans,
                                                  uint32_t id, int i, int j, uint32_t& lt, float (&q)[3], float& qs) {
#pragma clang fp contract(off)
  const uint32_t t = id >> 3;  // original index
  const int d = g_find_draw(p, t);
  const GDraw& dr = p.draws[d];
  lt = dr.inv[t - dr.tri_base];  // stored index
  const float* pp = dr.pos + (size_t)lt * 9;
  // the snapped vertices, orientation and area the raster used, from the
  // triangle's record (or its fan triangle's)
  const GRec& rec = recs[dr.tri_base + lt];
  const bool fan = (rec.swap & kClipped) != 0u;
  uint32_t fi = 0;  // fan record of a clipped triangle
  GSetup s;
  float cw[3];
  if (!fan) {
    s = g_setup_of(rec);
    // clip w of each vertex, as g_clip_coords computes it
    for (int v = 0; v < 3; ++v) {
      float w[3];
      world_of(dr, pp + 3 * v, w);
      float t4 = p.PV[3] * w[0];
      t4 = t4 + p.PV[7] * w[1];
      t4 = t4 + p.PV[11] * w[2];
      t4 = t4 + p.PV[15];
      cw[v] = t4;
    }
  } else {
    fi = (uint32_t)rec.area + (id & 7u);
    s = g_setup_of(fan_recs[fi]);
    for (int v = 0; v < 3; ++v) cw[v] = fans[fi].w[v];
  }
  float b[3];
  if (!g_cover(s, i, j, b[0], b[1], b[2])) {  // cannot happen: same arithmetic as the raster
    b[0] = b[1] = b[2] = 1.0f / 3.0f;
  }
  // perspective-correct weights: (b_i / w_i) / sum(b_k / w_k) over the drawn triangle
  qs = 0.0f;
  for (int v = 0; v < 3; ++v) {
    q[v] = b[v] / cw[v];
    qs = qs + q[v];
  }
  for (int v = 0; v < 3; ++v) q[v] = q[v] / qs;
  if (fan) {  // weights of the original vertices
    const GFan& fd = fans[fi];
    float lam[3];
    for (int o = 0; o < 3; ++o) {
      float a = q[0] * fd.wts[0][o];
      a = a + q[1] * fd.wts[1][o];
      a = a + q[2] * fd.wts[2][o];
      lam[o] = a;
    }
    for (int o = 0; o < 3; ++o) q[o] = lam[o];
  }
  return dr;
}

// Perspective-correct world position, normal, linear depth and material id of
// pixel (i, j) from its winning key; writes the G-buffer texel.
__device__ __forceinline__ void g_resolve_pixel(const GParams& p, const GRec* __restrict__ recs,
                                                const GRec* __restrict__ fan_recs, const GFan* __restrict__ fans,
                                                unsigned long long key, int i, int j, float4* __restrict__ gpos,
                                                float4* __restrict__ gnrm) {
#pragma clang fp contract(off)
  const size_t idx = (size_t)j * p.W + i;
  if (key == ~0ull) {
    gpos[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    gnrm[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  uint32_t lt;
  float q[3], qs;
  const GDraw& dr = g_weights(p, recs, fan_recs, fans, (uint32_t)(key & 0xffffffffu), i, j, lt, q, qs);
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

// 4. one workgroup per 32x32 tile: the depth test runs on a tile depth buffer
//    in LDS ((depth bits << 32 | original triangle id), 64-bit LDS atomicMin;
//    equal depths resolve to the lower triangle id, the draw-order winner of
//    GL_LESS).  The tile's triangles are taken kChunk at a time: their
//    records and edge functions at the first pixel of their box (clipped to
//    the tile) go to LDS, and the rows of all their boxes are dealt over the
//    lanes (a lane finds its triangle by a search over the prefix of the row
//    counts).  A lane walks its row's span with incremental edge functions,
//    so the work per lane is one row, not one triangle of any size.  Then
//    every pixel of the tile is resolved and written once.
constexpr int kChunk = 128;

// Row-walk data of one triangle of a chunk: edge values at (ci0, cj0) and
// their per-column / per-row increments, reciprocals of the per-column ones
// (0 when 0), the clipped box and the early-depth bound.
struct GRowSetup {
  int64_t w0[3], di[3], dj[3];
  int16_t ci0, ci1, cj0, pad;
};
static_assert(sizeof(GRowSetup) == 80, "GRowSetup must be 80 bytes");

// What else the tile kernel reads of a chunk's triangle, beside its rows:
// the barycentrics' divisor, the vertex depths and the depth key's triangle
// id (the exact phase), the span reciprocals and the early-depth bound (the
// walk).  40 B instead of the 64-B record: with the 80-B row data a chunk of
// 128 takes 15 KB, and the workgroup 25.5 KB of LDS, so 6 fit on a CU.
struct GTileRec {
  float fa;       // (float)area
  float z[3];     // window depth of each vertex
  uint32_t otri;  // the depth key's triangle id
  uint32_t swap;  // bit 0: vertices 1 and 2 swapped for CCW
  uint32_t zb;    // the early-depth bound, g_zmin_bits
  float inv[3];   // 1 / di per edge (0 when di == 0), for g_span
};
static_assert(sizeof(GTileRec) == 40, "GTileRec must be 40 bytes");

// Narrows [lo, hi] to the columns where w + d * (i - i0) >= 0 may hold, with
// inv_d = 1 / d computed once per triangle (0 when d == 0).  The crossing
// t = -w / d is bounded in float with a margin of 1e-3 pixel plus 2^-19 |t|
// (far above the rounding of the conversions, the reciprocal and the
// product), so it never drops a pixel the exact integer test would keep.
__device__ __forceinline__ void g_span(int64_t w, int64_t d, float inv_d, int i0, int& lo, int& hi) {
  if (d == 0) {
    if (w < 0) hi = lo - 1;  // the whole row is outside this edge
    return;
  }
  float t = (float)(-w) * inv_d;  // w(i) = 0 at i = i0 + t
  t = fminf(fmaxf(t, -4096.0f), 4096.0f);
  const float m = 1e-3f + fabsf(t) * 1.9073486e-6f;
  if (d > 0) lo = max(lo, i0 + (int)ceilf(t - m));
  else hi = min(hi, i0 + (int)floorf(t + m));
}


// Early depth rejection: a covered pixel's depth is a convex combination of
// the vertex depths, computed with b0 + b1 + b2 within 4 * 2^-23 of 1, so it
// is never below min(z) - 2e-6 (depths lie in [0, 1]).  A pixel whose tile
// depth is already below that bound keeps its winner whatever this triangle
// computes, so its exact barycentrics and atomic are skipped.  The key's high
// word is the depth's bit pattern (order-preserving for depths >= 0; the
// empty key ~0 never compares below).
__device__ __forceinline__ uint32_t g_zmin_bits(const GRec& r) {
  const float zmin = fminf(r.z[0], fminf(r.z[1], r.z[2])) - 2e-6f;
  return zmin > 0.0f ? __builtin_bit_cast(uint32_t, zmin) : 0u;
}

__device__ __forceinline__ bool g_hidden(const unsigned long long* s_depth, int li, int lj, uint32_t zmin_bits) {
  const uint32_t cur = (uint32_t)(s_depth[lj * kTile + li] >> 32);
  return zmin_bits > cur;
}

__device__ __forceinline__ void g_depth_test(unsigned long long* s_depth, int li, int lj, float b0, float b1,
                                             float b2, const GTileRec& r) {
#pragma clang fp contract(off)
  const float depth = b0 * r.z[0] + b1 * r.z[1] + b2 * r.z[2];
  if (depth >= 0.0f && depth <= 1.0f) {
    const unsigned long long key =
        ((unsigned long long)__builtin_bit_cast(uint32_t, depth) << 32) | (unsigned long long)r.otri;
    atomicMin(&s_depth[lj * kTile + li], key);
  }
}

// 6 waves per SIMD: the LDS of 6 workgroups fits a CU, and 80 registers
// hold the kernel without spills.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void g_tile_kernel(GParams p, const GRec* __restrict__ recs,
                                                     const GRec* __restrict__ fan_recs,
                                                     const GFan* __restrict__ fans,
                                                     const uint32_t* __restrict__ tile_count,
                                                     const uint32_t* __restrict__ tile_off,
                                                     const uint32_t* __restrict__ list, int tiles_x,
                                                     float4* __restrict__ gpos, float4* __restrict__ gnrm) {
#pragma clang fp contract(off)
  __shared__ unsigned long long s_depth[kTile * kTile];
  __shared__ GTileRec s_tr[kChunk];
  __shared__ GRowSetup s_row[kChunk];
  __shared__ uint32_t s_end[kChunk];  // inclusive prefix of the row counts
  __shared__ uint32_t s_wsum[kChunk / 64];
  __shared__ uint32_t s_q[4][128];  // per-wave queue of pixels that passed the edge signs and early depth test
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = blockIdx.x;
  const int x0 = (tile % tiles_x) * kTile, y0 = (tile / tiles_x) * kTile;
  for (int q = tid; q < kTile * kTile; q += 256) s_depth[q] = ~0ull;
  const uint32_t n = tile_count[tile], base = tile_off[tile];
  __syncthreads();
  for (uint32_t b0 = 0; b0 < n; b0 += kChunk) {
    if (tid < kChunk) {
      uint32_t rows = 0;
      if (b0 + tid < n) {
        const uint32_t e = list[base + b0 + tid];
        const GRec r = e < p.ntri ? recs[e] : fan_recs[e - p.ntri];
        const int ci0 = max((int)r.i0, x0), ci1 = min((int)r.i1, x0 + kTile - 1);
        const int cj0 = max((int)r.j0, y0), cj1 = min((int)r.j1, y0 + kTile - 1);
        if (ci1 >= ci0 && cj1 >= cj0) {
          rows = (uint32_t)(cj1 - cj0 + 1);
          GEdges e;
          g_edges(g_setup_of(r), ci0, cj0, e);
          GRowSetup g;
          GTileRec tr;
          for (int k = 0; k < 3; ++k) {
            g.w0[k] = e.w[k];
            g.di[k] = e.di[k];
            g.dj[k] = e.dj[k];
            tr.inv[k] = e.di[k] ? 1.0f / (float)e.di[k] : 0.0f;
            tr.z[k] = r.z[k];
          }
          g.ci0 = (int16_t)ci0;
          g.ci1 = (int16_t)ci1;
          g.cj0 = (int16_t)cj0;
          g.pad = 0;
          tr.fa = (float)r.area;
          tr.otri = r.otri;
          tr.swap = r.swap;
          tr.zb = g_zmin_bits(r);
          s_tr[tid] = tr;
          s_row[tid] = g;
        }
      }
      uint32_t x = rows;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane == 63) s_wsum[wave] = x;
      s_end[tid] = x;  // within the wave; the waves before are added below
    }
    __syncthreads();
    if (tid < kChunk) {
      uint32_t add = 0;
      for (int w = 0; w < wave; ++w) add += s_wsum[w];
      s_end[tid] += add;
    }
    __syncthreads();
    const uint32_t total = s_end[kChunk - 1];
    // Pixels in two phases.  1: each lane walks the span of its row with the
    // edge signs and the early depth rejection only, and appends the pixels
    // that pass to its wave's queue in LDS (ballot + popcount; no barrier:
    // only the wave touches it).  2: whenever the queue holds 64, every lane
    // takes one and runs the exact test (top-left rule), the barycentrics and
    // the depth test, so the expensive part runs with all lanes busy instead
    // of on the lanes whose rows are still long.  A pixel's key reaches its
    // atomic minimum whatever the order, so the result is the same.
    uint32_t* q = s_q[wave];
    uint32_t qn = 0;  // entries in the queue (the same in every lane)
    // queue entry: entry lo (7 bits) | row (5 bits) << 7 | column in the tile << 12
    auto pixel = [&](uint32_t d) {
      const uint32_t lo = d & 127u, row = (d >> 7) & 31u;
      const int li = (int)(d >> 12);
      const GRowSetup& g = s_row[lo];
      const GTileRec& tr = s_tr[lo];
      GEdges e;
      for (int k = 0; k < 3; ++k) {
        e.di[k] = g.di[k];
        e.ex[k] = g.dj[k] >> 8;   // dj = ex * 256, di = -ey * 256 (exact)
        e.ey[k] = -(g.di[k] >> 8);
      }
      const int64_t col = (int64_t)(x0 + li - (int)g.ci0);
      const int64_t wa = g.w0[0] + g.dj[0] * (int64_t)row + g.di[0] * col;
      const int64_t wb = g.w0[1] + g.dj[1] * (int64_t)row + g.di[1] * col;
      const int64_t wc = g.w0[2] + g.dj[2] * (int64_t)row + g.di[2] * col;
      float bw0, bw1, bw2;
      if (g_cover_w((tr.swap & 1u) != 0u, tr.fa, e, wa, wb, wc, bw0, bw1, bw2))
        g_depth_test(s_depth, li, (int)g.cj0 + (int)row - y0, bw0, bw1, bw2, tr);
    };
    uint32_t own0 = 0;  // the entry of the wave's first row (the same in every lane; never decreases)
    for (uint32_t c0 = (uint32_t)wave * 64u; c0 < total; c0 += 256) {
      // The entry (triangle) of each lane's row, for the window [c0, c0 + 64)
      // of consecutive rows.  own0 moves to the first entry ending after c0:
      // lane k tests entry own0 + k and the ballot counts those ending at or
      // before c0 (a prefix, the ends being sorted).  Lane k then holds the
      // end of entry own0 + k, and each lane finds how many of these end at
      // or before its row by a binary search over the lanes (register
      // shuffles): that count from own0 is its entry.  A window that spans
      // more than 64 entries finishes with a per-lane walk.
      const uint32_t c = c0 + (uint32_t)lane;
      for (;;) {
        const uint32_t ek = own0 + (uint32_t)lane < (uint32_t)kChunk ? s_end[own0 + lane] : 0xffffffffu;
        const uint32_t k = (uint32_t)__popcll(__ballot(ek <= c0));
        own0 += k;
        if (k < 64u) break;
      }
      const uint32_t ek = own0 + (uint32_t)lane < (uint32_t)kChunk ? s_end[own0 + lane] : 0xffffffffu;
      uint32_t past = 0;
      for (uint32_t st = 32; st; st >>= 1)
        if ((uint32_t)__shfl(ek, (int)(past + st - 1u), 64) <= c) past += st;
      if ((uint32_t)__shfl(ek, (int)past, 64) <= c) ++past;
      // the lane's row and its span: for each edge w(i) = w_row + di * (i - ci0)
      // >= 0, bounded in float with a margin (the exact tests decide)
      int len = 0, lo_i = 0, lj = 0;
      uint32_t lo = 0, row = 0, zb = 0;
      int64_t wa = 0, wb = 0, wc = 0, da = 0, db = 0, dc = 0;
      if (c < total) {
        lo = own0 + past;
        if (past == 64u)
          while (s_end[lo] <= c) ++lo;
        row = c - (lo ? s_end[lo - 1] : 0u);
        const GRowSetup& g = s_row[lo];
        const int ci0 = g.ci0, ci1 = g.ci1;
        lj = (int)g.cj0 + (int)row - y0;
        const int64_t ra = g.w0[0] + g.dj[0] * (int64_t)row;
        const int64_t rb = g.w0[1] + g.dj[1] * (int64_t)row;
        const int64_t rc = g.w0[2] + g.dj[2] * (int64_t)row;
        int hi_i = ci1;
        lo_i = ci0;
        const GTileRec& tr = s_tr[lo];
        g_span(ra, g.di[0], tr.inv[0], ci0, lo_i, hi_i);
        g_span(rb, g.di[1], tr.inv[1], ci0, lo_i, hi_i);
        g_span(rc, g.di[2], tr.inv[2], ci0, lo_i, hi_i);
        if (lo_i <= hi_i) {
          len = hi_i - lo_i + 1;
          const int64_t k0 = lo_i - ci0;
          da = g.di[0];
          db = g.di[1];
          dc = g.di[2];
          wa = ra + da * k0;
          wb = rb + db * k0;
          wc = rc + dc * k0;
          zb = tr.zb;
        }
      }
      // the wave walks its longest span; each lane stops at the end of its own
      for (int k = 0; __ballot(k < len) != 0ull; ++k) {
        bool hit = false;
        uint32_t desc = 0;
        if (k < len) {
          const int li = lo_i + k - x0;
          hit = wa >= 0 && wb >= 0 && wc >= 0 && !g_hidden(s_depth, li, lj, zb);
          desc = lo | (row << 7) | ((uint32_t)li << 12);
          wa += da;
          wb += db;
          wc += dc;
        }
        const unsigned long long m = __ballot(hit);
        if (hit) q[qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = desc;
        qn += (uint32_t)__popcll(m);
        if (qn >= 64u) {
          qn -= 64u;
          pixel(q[qn + (uint32_t)lane]);
        }
      }
    }
    if ((uint32_t)lane < qn) pixel(q[lane]);
    __syncthreads();
  }
  for (int q = tid; q < kTile * kTile; q += 256) {
    const int i = x0 + (q % kTile), j = y0 + (q / kTile);
    if (i < p.W && j < p.H) g_resolve_pixel(p, recs, fan_recs, fans, s_depth[q], i, j, gpos, gnrm);
  }
}

// ---- the voxel view: every fragment blended in draw order ---------------------
// Renderer::visualize with view_voxel_dir < 7 draws the models with GL_LESS
// depth and SRC_ALPHA / ONE_MINUS_SRC_ALPHA blending (renderer.cpp:355-390);
// the fragment shader returns textureLod(tex3D[dir], pos, lod)
// (voxel_cone_tracing.frag:253-257).  So a pixel shows, blended in draw order
// over the clear colour, every fragment nearer than all drawn before it.  The
// tile lists of the G-buffer raster are sorted into draw order (the records'
// triangle << 3 | fan key), then one workgroup per tile walks its triangles
// in that order; each lane owns 4 pixels of the tile (depth and colour in
// registers), so no ordering between lanes is needed.
struct VViewVol {
  const char* base;  // the volume (plain texels, Morton order per level)
  uint64_t off[16][6];
  int res, levels, dtype;
};

__device__ __forceinline__ float4 vv_texel(const VViewVol& V, int l, int map, int x, int y, int z) {
  const char* lv = V.base + V.off[l][l == 0 ? 0 : map];
  const uint32_t e = brick_index((uint32_t)x, (uint32_t)y, (uint32_t)z);
  if (V.dtype == VCT_RGBA8) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(lv)[e];
    return make_float4(unorm8_exact(v & 255u), unorm8_exact((v >> 8) & 255u), unorm8_exact((v >> 16) & 255u),
                       unorm8_exact(v >> 24));
  }
  if (V.dtype == VCT_FP16 || V.dtype == VCT_BF16) {
    const uint2 v = reinterpret_cast<const uint2*>(lv)[e];
    if (V.dtype == VCT_FP16)
      return make_float4(f16_to_f32((uint16_t)(v.x & 0xffffu)), f16_to_f32((uint16_t)(v.x >> 16)),
                         f16_to_f32((uint16_t)(v.y & 0xffffu)), f16_to_f32((uint16_t)(v.y >> 16)));
    return make_float4(bf16_to_f32((uint16_t)(v.x & 0xffffu)), bf16_to_f32((uint16_t)(v.x >> 16)),
                       bf16_to_f32((uint16_t)(v.y & 0xffffu)), bf16_to_f32((uint16_t)(v.y >> 16)));
  }
  return reinterpret_cast<const float4*>(lv)[e];
}

// One level, trilinear with a zero border (CLAMP_TO_BORDER, texture_3d.cpp:10-15).
__device__ __forceinline__ float4 vv_trilinear(const VViewVol& V, int l, int map, float px, float py, float pz) {
#pragma clang fp contract(off)
  const int s = V.res >> l;
  const float u = px * (float)s - 0.5f, v = py * (float)s - 0.5f, w = pz * (float)s - 0.5f;
  const float fu = floorf(u), fv = floorf(v), fw = floorf(w);
  const int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
  const float a = u - fu, b = v - fv, c = w - fw;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int x = i0 + dx, y = j0 + dy, z = k0 + dz;
        if (x < 0 || y < 0 || z < 0 || x >= s || y >= s || z >= s) continue;
        const float wt = (dx ? a : 1.0f - a) * (dy ? b : 1.0f - b) * (dz ? c : 1.0f - c);
        const float4 t = vv_texel(V, l, map, x, y, z);
        out.x = out.x + wt * t.x;
        out.y = out.y + wt * t.y;
        out.z = out.z + wt * t.z;
        out.w = out.w + wt * t.w;
      }
  return out;
}

// textureLod with LINEAR_MIPMAP_LINEAR, lod clamped to [0, levels - 1].
__device__ __forceinline__ float4 vv_texture_lod(const VViewVol& V, int map, float px, float py, float pz,
                                                 float lod) {
#pragma clang fp contract(off)
  const float lmax = (float)(V.levels - 1);
  if (!(lod > 0.0f)) lod = 0.0f;
  if (lod > lmax) lod = lmax;
  const int d1 = (int)floorf(lod);
  const float f = lod - (float)d1;
  const float4 a = vv_trilinear(V, d1, map, px, py, pz);
  if (!(f > 0.0f && d1 + 1 < V.levels)) return a;
  const float4 b = vv_trilinear(V, d1 + 1, map, px, py, pz);
  const float g = 1.0f - f;
  return make_float4(g * a.x + f * b.x, g * a.y + f * b.y, g * a.z + f * b.z, g * a.w + f * b.w);
}

// Keys (the record's triangle << 3 | fan) and values (list entries) of every
// tile's list, and the end of each tile's segment, for the sort.
__global__ __launch_bounds__(256) void g_vview_keys_kernel(const uint32_t* __restrict__ tile_count,
                                                           const uint32_t* __restrict__ tile_off,
                                                           const uint32_t* __restrict__ list,
                                                           const GRec* __restrict__ recs,
                                                           const GRec* __restrict__ fan_recs, uint32_t ntri,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ ends) {
  const uint32_t tile = blockIdx.x, n = tile_count[tile], base = tile_off[tile];
  if (threadIdx.x == 0) ends[tile] = base + n;
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    const uint32_t e = list[base + k];
    keys[base + k] = (e < ntri ? recs[e] : fan_recs[e - ntri]).otri;
    vals[base + k] = e;
  }
}

struct VViewArgs {
  GParams p;
  const GRec* recs;
  const GRec* fan_recs;
  const GFan* fans;
  const uint32_t* tile_count;
  const uint32_t* tile_off;
  const uint32_t* order;  // each tile's list entries in draw order
  int tiles_x;
  VViewVol V;
  int map;
  float lod, cube_size;
  float4* image;
};

__global__ __launch_bounds__(256) void g_vview_kernel(VViewArgs A) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  const int tile = blockIdx.x;
  const int x0 = (tile % A.tiles_x) * kTile, y0 = (tile / A.tiles_x) * kTile;
  constexpr int kPer = kTile * kTile / 256;  // pixels per lane
  float dep[kPer];
  float4 col[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    dep[q] = __builtin_inff();
    col[q] = make_float4(0.15f, 0.25f, 0.25f, 1.0f);  // clear colour, renderer.cpp:398
  }
  const uint32_t n = A.tile_count[tile], base = A.tile_off[tile];
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t e = A.order[base + k];
    const GRec r = e < A.p.ntri ? A.recs[e] : A.fan_recs[e - A.p.ntri];
    if (r.i0 > r.i1) continue;
    const GSetup s = g_setup_of(r);
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int li = (tid + 256 * q) % kTile, lj = (tid + 256 * q) / kTile;
      const int i = x0 + li, j = y0 + lj;
      if (i < r.i0 || i > r.i1 || j < r.j0 || j > r.j1) continue;
      float b0, b1, b2;
      if (!g_cover(s, i, j, b0, b1, b2)) continue;
      const float depth = b0 * r.z[0] + b1 * r.z[1] + b2 * r.z[2];
      if (!(depth >= 0.0f && depth <= 1.0f) || !(depth < dep[q])) continue;  // GL_LESS
      dep[q] = depth;
      uint32_t lt;
      float w[3], qs;
      const GDraw& dr = g_weights(A.p, A.recs, A.fan_recs, A.fans, r.otri, i, j, lt, w, qs);
      const float* pp = dr.pos + (size_t)lt * 9;
      float wp[3] = {0.f, 0.f, 0.f};
      for (int v = 0; v < 3; ++v) {
        float wv[3];
        world_of(dr, pp + 3 * v, wv);
        for (int c = 0; c < 3; ++c) wp[c] = wp[c] + w[v] * wv[c];
      }
      const float bx = wp[0] / A.cube_size * 0.5f + 0.5f, by = wp[1] / A.cube_size * 0.5f + 0.5f,
                  bz = wp[2] / A.cube_size * 0.5f + 0.5f;
      // within_cube (:250): outside, the fragment writes its depth and no colour
      if (!(fabsf(bx) < 1.0f && fabsf(by) < 1.0f && fabsf(bz) < 1.0f)) continue;
      const float4 sm = vv_texture_lod(A.V, A.map, bx, by, bz, A.lod);
      const float a = sm.w, g = 1.0f - a;
      col[q] = make_float4(sm.x * a + col[q].x * g, sm.y * a + col[q].y * g, sm.z * a + col[q].z * g,
                           a * a + col[q].w * g);
    }
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int i = x0 + (tid + 256 * q) % kTile, j = y0 + (tid + 256 * q) / kTile;
    if (i < A.p.W && j < A.p.H) A.image[(size_t)j * A.p.W + i] = col[q];
  }
}

// workgroups of the kernels that stride over the clipped triangles
constexpr int kFanGrid = 64;

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
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

int vct_launch_gbuffer(vct_context* ctx) { return vct_launch_gbuffer_phase(ctx, 3); }

int vct_launch_gbuffer_phase(vct_context* ctx, int phase) {
  const int W = ctx->width, H = ctx->height;
  if (W > 32767 || H > 32767) return vct_set_error(VCT_ERR_LIMIT, "viewport above 32767 pixels");
  const size_t npix = (size_t)W * H;
  int rc;
  if ((rc = vct_ensure(ctx->g_pos, npix * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->g_nrm, npix * 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->counters, 256, ctx))) return rc;
  uint32_t* counters = (uint32_t*)ctx->counters.p;

  GParams p;
  memset(&p, 0, sizeof(p));
  uint64_t ntri = 0;
  mat4_mul(ctx->proj, ctx->view, p.PV);
  // every queued model is one draw, in queue order (renderer.cpp:242-256)
  std::vector<GDraw> draws;
  draws.reserve(ctx->queue.size());
  for (uint32_t id : ctx->queue) {
    if (id >= ctx->models.size() || !ctx->models[id].ntri) continue;
    const VctModel& m = ctx->models[id];
    draws.emplace_back();
    GDraw& d = draws.back();
    memset(&d, 0, sizeof(d));
    d.pos = m.d_pos;
    d.nrm = m.d_nrm;
    d.mat = m.d_mat;
    d.perm = m.d_perm;
    d.inv = m.d_inv;
    d.tri_base = (uint32_t)ntri;
    d.ntri = m.ntri;
    memcpy(d.M, m.xform, 64);
    normal_matrix3(m.xform, d.N);
    ntri += m.ntri;
    // depth keys and tile-list entries hold triangle << 3 | fan triangle
    if (ntri >= (1ull << 29)) return vct_set_error(VCT_ERR_LIMIT, "at most 2^29 triangles per frame");
  }
  p.ndraw = (int)draws.size();
  if ((phase & 1) && p.ndraw) {
    if ((rc = vct_upload(ctx, ctx->g_draws, draws.data(), draws.size() * sizeof(GDraw)))) return rc;
  }
  p.draws = (const GDraw*)ctx->g_draws.dev.p;
  p.ntri = (uint32_t)ntri;
  p.W = W;
  p.H = H;
  const int tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + kTile - 1) / kTile;
  const uint32_t ntiles = (uint32_t)(tiles_x * tiles_y);
  // [0, ntiles): counts, [ntiles, 2 ntiles): write cursors
  if ((rc = vct_ensure(ctx->g_count, (size_t)ntiles * 8 + 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->g_off, (size_t)ntiles * 4 + 16, ctx))) return rc;
  uint32_t* tile_count = (uint32_t*)ctx->g_count.p;
  uint32_t* cursor = tile_count + ntiles;
  uint32_t* tile_off = (uint32_t*)ctx->g_off.p;
  if (phase & 1) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3((2 * ntiles + 255) / 256), dim3(256), 0, ctx->stream, tile_count,
                       2 * ntiles, 0u);
    VCT_HIP(hipGetLastError());
    ctx->h_readback[0] = 0u;
    ctx->h_readback[1] = 0u;
    if (ntri) {
      if ((rc = vct_ensure(ctx->g_recs, (size_t)ntri * sizeof(GRec), ctx))) return rc;
      if ((rc = vct_ensure(ctx->g_clip, (size_t)ntri * 4, ctx))) return rc;
      hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, ctx->stream, counters + 5, 2u, 0u);
      hipLaunchKernelGGL(g_bin_count_kernel, dim3((ntri + 255) / 256), dim3(256), 0, ctx->stream, p,
                         (GRec*)ctx->g_recs.p, tile_count, tiles_x, (uint32_t*)ctx->g_clip.p, counters + 6);
      hipLaunchKernelGGL(g_fan_count_kernel, dim3(kFanGrid), dim3(256), 0, ctx->stream, p, (GRec*)ctx->g_recs.p,
                         (const uint32_t*)ctx->g_clip.p, (const uint32_t*)(counters + 6), tile_count, tiles_x,
                         counters + 5);
      VCT_HIP(hipGetLastError());
      if ((rc = vct_launch_scan_u32(ctx, tile_count, tile_off, ntiles, counters + 4, &ctx->g_scan_tmp))) return rc;
      // the list size and the number of fan triangles of clipped triangles
      // are known on the device only
      VCT_HIP(hipMemcpyAsync(ctx->h_readback, counters + 4, 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    // also when nothing was counted: a copy an interrupted earlier frame left
    // in flight must not land after the words are read
    VCT_HIP(hipEventRecord(ctx->g_count_ev, ctx->stream));
  }
  if (!(phase & 2)) return VCT_OK;
  VCT_HIP(hipEventSynchronize(ctx->g_count_ev));
  const uint32_t total = ctx->h_readback[0];
  const uint32_t nfans = ctx->h_readback[1];
  if ((rc = vct_ensure(ctx->g_fans, (size_t)nfans * (sizeof(GRec) + sizeof(GFan)) + 64, ctx))) return rc;
  GRec* fan_recs = (GRec*)ctx->g_fans.p;
  GFan* fans = (GFan*)((char*)ctx->g_fans.p + (size_t)nfans * sizeof(GRec));
  if (total) {
    if ((rc = vct_ensure(ctx->g_list, (size_t)total * 4 + (size_t)total / 4 * 4, ctx))) return rc;
    hipLaunchKernelGGL(g_bin_write_kernel, dim3((p.ntri + 255) / 256), dim3(256), 0, ctx->stream, p.ntri,
                       (const GRec*)ctx->g_recs.p, (const uint32_t*)tile_off, cursor, tiles_x,
                       (uint32_t*)ctx->g_list.p);
    if (nfans)
      hipLaunchKernelGGL(g_fan_write_kernel, dim3(kFanGrid), dim3(256), 0, ctx->stream, p, (const GRec*)ctx->g_recs.p,
                         (const uint32_t*)ctx->g_clip.p, (const uint32_t*)(counters + 6), (const uint32_t*)tile_off,
                         cursor, tiles_x, (uint32_t*)ctx->g_list.p, fan_recs, fans);
    VCT_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL(fill_u32_kernel, dim3((ntiles + 255) / 256), dim3(256), 0, ctx->stream, tile_off, ntiles,
                       0u);
    VCT_HIP(hipGetLastError());
    if ((rc = vct_ensure(ctx->g_list, 16, ctx))) return rc;
  }
  if ((rc = vct_ensure(ctx->g_recs, 64, ctx))) return rc;
  hipLaunchKernelGGL(g_tile_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, p, (const GRec*)ctx->g_recs.p,
                     (const GRec*)fan_recs, (const GFan*)fans, (const uint32_t*)tile_count, (const uint32_t*)tile_off, (const uint32_t*)ctx->g_list.p,
                     tiles_x, (float4*)ctx->g_pos.p, (float4*)ctx->g_nrm.p);
  VCT_HIP(hipGetLastError());
  ctx->last_gbuffer_refs = total;
  ctx->g_frame.valid = true;
  ctx->g_frame.W = W;
  ctx->g_frame.H = H;
  ctx->g_frame.ndraw = p.ndraw;
  ctx->g_frame.ntri = p.ntri;
  ctx->g_frame.nfans = nfans;
  ctx->g_frame.total = total;
  memcpy(ctx->g_frame.PV, p.PV, sizeof(p.PV));
  return VCT_OK;
}

// The voxel view of the last rasterised G-buffer, blended in draw order: the
// tile lists sorted by draw order (a segmented radix sort, one segment per
// tile), then one workgroup per tile.
int vct_launch_voxel_view(vct_context* ctx) {
  const auto& F = ctx->g_frame;
  const int W = F.W, H = F.H;
  const int tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + kTile - 1) / kTile;
  const uint32_t ntiles = (uint32_t)(tiles_x * tiles_y);
  const uint64_t total = F.total;
  if (total > 0xffffffffull) return vct_set_error(VCT_ERR_LIMIT, "voxel view: more than 2^32 - 1 tile entries");
  int rc;
  const size_t n4 = (size_t)(total ? total : 1) * 4;
  if ((rc = vct_ensure(ctx->vv_keys, 2 * n4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->vv_vals, 2 * n4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->vv_ends, (size_t)ntiles * 4, ctx))) return rc;
  uint32_t* keys = (uint32_t*)ctx->vv_keys.p;
  uint32_t* vals = (uint32_t*)ctx->vv_vals.p;
  const GRec* fan_recs = (const GRec*)ctx->g_fans.p;
  const GFan* fans = (const GFan*)((const char*)ctx->g_fans.p + (size_t)F.nfans * sizeof(GRec));
  hipLaunchKernelGGL(g_vview_keys_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, (const uint32_t*)ctx->g_count.p,
                     (const uint32_t*)ctx->g_off.p, (const uint32_t*)ctx->g_list.p, (const GRec*)ctx->g_recs.p,
                     fan_recs, F.ntri, keys, vals, (uint32_t*)ctx->vv_ends.p);
  VCT_HIP(hipGetLastError());
  size_t temp = 0;
  VCT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, temp, keys, keys + total, vals, vals + total,
                                                       (int)total, (int)ntiles, (const uint32_t*)ctx->g_off.p,
                                                       (const uint32_t*)ctx->vv_ends.p, 0, 32, ctx->stream));
  if ((rc = vct_ensure(ctx->vv_temp, temp ? temp : 16, ctx))) return rc;
  if (total)
    VCT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(ctx->vv_temp.p, temp, keys, keys + total, vals, vals + total,
                                                         (int)total, (int)ntiles, (const uint32_t*)ctx->g_off.p,
                                                         (const uint32_t*)ctx->vv_ends.p, 0, 32, ctx->stream));
  VViewArgs A;
  memset(&A, 0, sizeof(A));
  A.p.draws = (const GDraw*)ctx->g_draws.dev.p;
  A.p.ndraw = F.ndraw;
  memcpy(A.p.PV, F.PV, sizeof(A.p.PV));
  A.p.ntri = F.ntri;
  A.p.W = W;
  A.p.H = H;
  A.recs = (const GRec*)ctx->g_recs.p;
  A.fan_recs = fan_recs;
  A.fans = fans;
  A.tile_count = (const uint32_t*)ctx->g_count.p;
  A.tile_off = (const uint32_t*)ctx->g_off.p;
  A.order = vals + total;
  A.tiles_x = tiles_x;
  A.V.base = (const char*)ctx->volume.p;
  for (int l = 0; l < ctx->levels && l < 16; ++l)
    for (int d = 0; d < 6; ++d) A.V.off[l][d] = ctx->lvl_off[l][d];
  A.V.res = (int)ctx->res;
  A.V.levels = ctx->levels;
  A.V.dtype = ctx->dtype;
  A.map = ctx->view_voxel_dir < 0 ? 0 : (ctx->view_voxel_dir > 5 ? 5 : ctx->view_voxel_dir);
  A.lod = ctx->view_voxel_lod;
  A.cube_size = ctx->cube_size;
  A.image = (float4*)ctx->image.p;
  hipLaunchKernelGGL(g_vview_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, A);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}
