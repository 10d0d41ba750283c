// vct_voxelize.hip - voxelisation (kernel a).
//
// Replaces the voxelize program: voxelize.vert:24-30 (world position divided
// by cube_size, normal by the inverse-transpose model matrix),
// voxelize.geom:25-55 (projection onto the plane of the dominant axis of
// |cross(e1, e2)|, ties to XZ), GL rasterisation at pixel centres of a
// 2res x 2res viewport (renderer.cpp:339-343), and voxelize.frag:122-161
// (per-fragment direct lighting and the collision average into the voxel at
// ivec3(res * (0.5 p + 0.5))).
//
// MI355X design
//  1. setup (one lane per triangle): dominant axis, vertices snapped to
//     1/256 pixel, pixel bounding box in 2x2 quads, a 128-byte record; each
//     triangle is counted into the 8^3-voxel bricks its voxel box overlaps.
//  2. exclusive scan of the brick counts, then per-brick triangle lists.
//  3. one workgroup per brick: the pixels of its triangles' boxes inside the
//     brick's footprint are spread over the lanes, coverage is tested with
//     exact 64-bit edge functions (top-left rule) from records in LDS, and
//     only fragments whose voxel lies in the brick are shaded, so every
//     fragment is found exactly once.
//  4a. exact mean (any storage): 16.16 fixed-point sums in LDS (exact, so
//     order-independent); the workgroup writes its brick's texels once.
//  4b. reference running average (RGBA8): a count pass sizes every voxel's
//     slots (scan over the voxels), a place pass stores (canonical candidate
//     key, colour) in them and then sorts each voxel's slots and replays the
//     packed RGBA8 + count-nibble recurrence of imageAtomicRGBA8Avg
//     (voxelize.frag:95-120) - the result no longer depends on scheduling.
//     No counter is shared between workgroups (a single append counter for
//     all fragments serialises at one L2 channel).
// Bit-exact parts (coverage, voxel coordinates, fragment colours) use no
// fused multiply-add and correctly rounded division / square root.
#include <cmath>
#include <cstring>

#include <algorithm>

#include "vct_internal.h"

namespace vct {

struct Draw {
  const float* pos;
  const float* nrm;
  const uint32_t* mat;
  const uint32_t* perm;  // stored -> original triangle index
  uint32_t tri_base;    // first triangle of this draw in the batch
  uint32_t canon_base;  // original index of the draw's first triangle in the frame
  uint32_t ntri;
  float M[16];  // model matrix, column-major
  float N[9];   // mat3(transpose(inverse(model))), column-major
};

// The queued draws of a frame, in queue order, in device memory (no limit on
// their number).
struct DrawTable {
  const Draw* d;
  int ndraw;
  uint32_t ntri;
};

struct VoxParams {
  uint32_t res;
  float inv_unused;
  float cube_size;
  int average;   // VCT_AVG_*
  uint32_t unused;
  LightSet lights;
  // normalize(position) of each directional light, computed on the host with
  // the same IEEE float operations (sqrt, then three divisions), so the bits
  // equal what every fragment would compute
  float ldir[VCT_MAX_LIGHTS][3];
};

struct Setup {
  float w[3][3];   // world / cube_size
  int64_t X[3], Y[3];
  int axis;
  int qx0, qy0, qw, qh;
  int64_t area;
  bool swap;
  bool ok;
};

// Per-triangle record written once by the setup pass and read by the brick
// workgroups (coverage fields through LDS, the rest for covered pixels).
struct alignas(16) TriRec {
  int64_t area;     // doubled signed area in 1/256 pixel units, made positive
  int32_t X[3], Y[3];
  int32_t qx0, qy0, qw;
  uint32_t mat;
  uint32_t flags;   // bit 0: vertices 1, 2 swapped (CW); bits 1-2: dominant axis; bits 8-31: qh
  float w[3][3];    // world position / cube_size per vertex
  float n[3][3];    // normalised transformed normal per vertex
  uint32_t otri;   // original triangle index in the frame (canonical order)
};
static_assert(sizeof(TriRec) == 128, "TriRec must be 128 bytes");

// The material terms a fragment of the triangle needs (voxelize.frag:139-155),
// gathered per triangle by the set-up kernel in frame order, so a fragment
// reads them by triangle index rather than through its record's material id
// (a dependent gather into the material table; on a soup of 1M triangles
// with a material each the brick pass went from 360 to 336 us).  They depend
// on the materials and the queued models only, so frames that change
// neither keep them.
struct alignas(16) VoxShade {
  float diffuse[3], alpha;   // alpha: the dissolve of a transmissive material, else 1
  float emission[3], pad0;
  float tr[3], pad1;         // the transmittance of a transmissive material, else 1 (1 * c == c)
};
static_assert(sizeof(VoxShade) == 48, "VoxShade must be 48 bytes");

__device__ __forceinline__ VoxShade make_shade(const DevMaterial& m) {
  VoxShade h;
  const bool t = illum_transmissive(m.illum);
  for (int k = 0; k < 3; ++k) {
    h.diffuse[k] = m.diffuse[k];
    h.emission[k] = m.emission[k];
    h.tr[k] = t ? m.transmittance[k] : 1.0f;
  }
  h.alpha = t ? m.dissolve : 1.0f;
  h.pad0 = h.pad1 = 0.0f;
  return h;
}

// The draw holding stored triangle t: the last one with tri_base <= t.
__device__ __forceinline__ int find_draw(const DrawTable& b, uint32_t t) {
  int lo = 0, hi = b.ndraw - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.d[mid].tri_base <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// voxelize.vert:26 + voxelize.geom:27-51 + viewport + snap.
__device__ __forceinline__ void tri_setup(const Draw& dr, uint32_t lt, uint32_t res, float cube_size, Setup& s) {
#pragma clang fp contract(off)
  const float* p = dr.pos + (size_t)lt * 9;
  for (int v = 0; v < 3; ++v) {
    const float px = p[v * 3 + 0], py = p[v * 3 + 1], pz = p[v * 3 + 2];
    for (int k = 0; k < 3; ++k) {
      float t = dr.M[0 + k] * px;
      t = t + dr.M[4 + k] * py;
      t = t + dr.M[8 + k] * pz;
      t = t + dr.M[12 + k];
      s.w[v][k] = t / cube_size;
    }
  }
  float e1[3], e2[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = s.w[1][k] - s.w[0][k];
    e2[k] = s.w[2][k] - s.w[0][k];
  }
  float cx = fabsf(e1[1] * e2[2] - e1[2] * e2[1]);
  float cy = fabsf(e1[2] * e2[0] - e1[0] * e2[2]);
  float cz = fabsf(e1[0] * e2[1] - e1[1] * e2[0]);
  int axis, ia, ib;
  if (cz > cx && cz > cy) {
    axis = 2; ia = 0; ib = 1;   // XY plane
  } else if (cx > cy && cx > cz) {
    axis = 0; ia = 1; ib = 2;   // YZ plane
  } else {
    axis = 1; ia = 0; ib = 2;   // XZ plane
  }
  s.axis = axis;
  const float fres = (float)res;
  bool ok = true;
  for (int v = 0; v < 3; ++v) {
    float a = s.w[v][ia], b = s.w[v][ib];
    if (!(fabsf(a) < 64.0f && fabsf(b) < 64.0f)) ok = false;
    float xw = a * fres + fres;  // viewport transform, scale = translate = res
    float yw = b * fres + fres;
    s.X[v] = (int64_t)__builtin_rintf((xw - 0.5f) * 256.0f);
    s.Y[v] = (int64_t)__builtin_rintf((yw - 0.5f) * 256.0f);
  }
  int64_t area = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
  s.swap = area < 0;
  s.area = area < 0 ? -area : area;
  if (area == 0) ok = false;
  int64_t xmin = min(s.X[0], min(s.X[1], s.X[2])), xmax = max(s.X[0], max(s.X[1], s.X[2]));
  int64_t ymin = min(s.Y[0], min(s.Y[1], s.Y[2])), ymax = max(s.Y[0], max(s.Y[1], s.Y[2]));
  const int64_t lim = 2 * (int64_t)res - 1;
  int64_t imin = (xmin + 255) >> 8, imax = xmax >> 8, jmin = (ymin + 255) >> 8, jmax = ymax >> 8;
  imin = imin < 0 ? 0 : imin;
  jmin = jmin < 0 ? 0 : jmin;
  imax = imax > lim ? lim : imax;
  jmax = jmax > lim ? lim : jmax;
  if (imin > imax || jmin > jmax) ok = false;
  s.ok = ok;
  if (ok) {
    s.qx0 = (int)(imin & ~1ll);
    s.qy0 = (int)(jmin & ~1ll);
    s.qw = (int)((imax - s.qx0) >> 1) + 1;
    s.qh = (int)((jmax - s.qy0) >> 1) + 1;
  } else {
    s.qx0 = s.qy0 = s.qw = s.qh = 0;
  }
}

// Record of a triangle that passed setup (s.ok).
__device__ __forceinline__ TriRec make_rec(const Draw& dr, uint32_t lt, const Setup& s) {
#pragma clang fp contract(off)
  TriRec r;
  r.area = s.area;
  for (int v = 0; v < 3; ++v) {
    r.X[v] = (int32_t)s.X[v];
    r.Y[v] = (int32_t)s.Y[v];
    for (int k = 0; k < 3; ++k) r.w[v][k] = s.w[v][k];
  }
  r.qx0 = s.qx0;
  r.qy0 = s.qy0;
  r.qw = s.qw;
  r.mat = dr.mat[lt];
  r.flags = (s.swap ? 1u : 0u) | ((uint32_t)s.axis << 1) | ((uint32_t)s.qh << 8);
  r.otri = dr.canon_base + dr.perm[lt];
  // normal: voxelize.vert:28 per vertex (normalize(N * n))
  const float* nr = dr.nrm + (size_t)lt * 9;
  for (int v = 0; v < 3; ++v) {
    float tt[3];
    for (int k = 0; k < 3; ++k) {
      float a = dr.N[0 + k] * nr[v * 3 + 0];
      a = a + dr.N[3 + k] * nr[v * 3 + 1];
      a = a + dr.N[6 + k] * nr[v * 3 + 2];
      tt[k] = a;
    }
    const float l = sqrtf(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
    for (int k = 0; k < 3; ++k) r.n[v][k] = tt[k] / l;
  }
  return r;
}

__device__ __forceinline__ bool top_left(int64_t dx, int64_t dy) { return dy < 0 || (dy == 0 && dx < 0); }

// Edge functions of a record in CCW order (a, b, c) = (0, 1, 2) or (0, 2, 1)
// at pixel (i, j), and their exact per-column / per-row increments.
struct VEdges {
  int64_t w[3];
  int64_t ex[3], ey[3];
};

// Differences of snapped coordinates fit 32 bits (vertices lie within 65
// grid widths of the origin: |X|, |Y| < 2^25 in 1/256 pixel, and 256 i,
// 256 j < 2^20), so every product is a 32 x 32 -> 64-bit multiply.
template <typename Rec>
__device__ __forceinline__ void vox_edges(const Rec& s, int i, int j, VEdges& e) {
  const int32_t px = i * 256, py = j * 256;
  const bool sw = (s.flags & 1u) != 0u;
  const int32_t ax = s.X[0], ay = s.Y[0];
  const int32_t bx = sw ? s.X[2] : s.X[1], by = sw ? s.Y[2] : s.Y[1];
  const int32_t cx = sw ? s.X[1] : s.X[2], cy = sw ? s.Y[1] : s.Y[2];
  e.w[0] = (int64_t)(cx - bx) * (int64_t)(py - by) - (int64_t)(cy - by) * (int64_t)(px - bx);
  e.w[1] = (int64_t)(ax - cx) * (int64_t)(py - cy) - (int64_t)(ay - cy) * (int64_t)(px - cx);
  e.w[2] = (int64_t)(bx - ax) * (int64_t)(py - ay) - (int64_t)(by - ay) * (int64_t)(px - ax);
  e.ex[0] = cx - bx;
  e.ey[0] = cy - by;
  e.ex[1] = ax - cx;
  e.ey[1] = ay - cy;
  e.ex[2] = bx - ax;
  e.ey[2] = by - ay;
}

// Coverage of a pixel from its edge values (top-left rule).
__device__ __forceinline__ bool vox_covers(const VEdges& e) {
  if (e.w[0] < 0 || e.w[1] < 0 || e.w[2] < 0) return false;
  if (e.w[0] == 0 && !top_left(e.ex[0], e.ey[0])) return false;
  if (e.w[1] == 0 && !top_left(e.ex[1], e.ey[1])) return false;
  if (e.w[2] == 0 && !top_left(e.ex[2], e.ey[2])) return false;
  return true;
}

// Fragment at pixel (i, j) of the triangle's 2res x 2res raster given its
// edge values (wa, wb, wc): coverage, voxel (Morton index) and colour.
// Returns false if the pixel is not covered or the voxel falls outside the
// grid (with kBrick: also when the voxel's Morton index >> bshift is not
// bwant, before any shading).  kKeyOnly stops once the voxel is known.
// `c` holds the coverage fields (X, Y, area, flags), `sp` the whole record
// and `sh` the triangle's material terms (both read only for covered pixels).
template <bool kBrick = false, bool kKeyOnly = false, typename Rec>
__device__ __forceinline__ bool shade_w(const Rec& c, const TriRec* sp, const VoxShade* sh, const VEdges& e, int i,
                                        int j, int64_t wa, int64_t wb, int64_t wc, const VoxParams& vp,
                                        uint32_t& key, float4& color, uint32_t bshift = 0, uint32_t bwant = 0) {
#pragma clang fp contract(off)
  const int lim = 2 * (int)vp.res;
  if (i >= lim || j >= lim) return false;
  const bool sw = (c.flags & 1u) != 0u;
  if (wa < 0 || wb < 0 || wc < 0) return false;
  if (wa == 0 && !top_left(e.ex[0], e.ey[0])) return false;
  if (wb == 0 && !top_left(e.ex[1], e.ey[1])) return false;
  if (wc == 0 && !top_left(e.ex[2], e.ey[2])) return false;
  // weights of the original vertices 0, 1, 2
  const int64_t w0 = wa, w1 = sw ? wc : wb, w2 = sw ? wb : wc;
  const float fa = (float)c.area;
  const float b0 = (float)w0 / fa, b1 = (float)w1 / fa, b2 = (float)w2 / fa;
  const TriRec& s = *sp;
  float pos[3];
  for (int k = 0; k < 3; ++k) pos[k] = b0 * s.w[0][k] + b1 * s.w[1][k] + b2 * s.w[2][k];
  // voxel address, voxelize.frag:156-157 (truncation toward zero)
  int v[3];
  const float fres = (float)vp.res;
  for (int k = 0; k < 3; ++k) {
    float f = fres * (0.5f * pos[k] + 0.5f);
    if (!(f > -1.0f && f < fres)) return false;
    v[k] = (int)f;
  }
  key = brick_index((uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2]);
  if (kBrick && (key >> bshift) != bwant) return false;
  if (kKeyOnly) return true;
  // normal: per-vertex normals (record), interpolated, normalised in the
  // fragment shader (voxelize.frag:134)
  float n[3];
  for (int k = 0; k < 3; ++k) n[k] = b0 * s.n[0][k] + b1 * s.n[1][k] + b2 * s.n[2][k];
  {
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] = n[k] / l;
  }
  // lighting, voxelize.frag:127-143
  float col[3] = {0.0f, 0.0f, 0.0f};
  const int nl = vp.lights.count < VCT_MAX_LIGHTS ? vp.lights.count : VCT_MAX_LIGHTS;
  for (int li = 0; li < nl; ++li) {
    const DevLight& L = vp.lights.l[li];
    float dir[3];
    float att;
    if (L.kind == VCT_LIGHT_DIRECTIONAL) {
      for (int k = 0; k < 3; ++k) dir[k] = vp.ldir[li][k];
      att = 1.0f;
    } else {
      float dv[3];
      for (int k = 0; k < 3; ++k) dv[k] = L.pos[k] / vp.cube_size - pos[k];
      float d = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      for (int k = 0; k < 3; ++k) dir[k] = dv[k] / d;
      att = 1.0f / (1.0f + d * d);
    }
    float c = n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2];
    c = c > 0.0f ? c : 0.0f;
    float ca = c * att;
    for (int k = 0; k < 3; ++k) col[k] = col[k] + ca * L.color[k] * L.intensity;
  }
  const VoxShade m = *sh;
  float out[4];
  for (int k = 0; k < 3; ++k) {
    float c = m.diffuse[k] * col[k] + m.emission[k];
    out[k] = m.tr[k] * c;
  }
  out[3] = m.alpha;
  for (int k = 0; k < 4; ++k) out[k] = fminf(fmaxf(out[k], 0.0f), 1.0f);
  color = make_float4(out[0], out[1], out[2], out[3]);
  return true;
}

// Buffers of the reference running average's two brick passes.
struct RefPass {
  uint32_t* vcount;              // count pass: fragments per voxel; place pass: fragments placed so far
  const uint32_t* vox_off;       // first slot of each voxel's fragments (exclusive scan of the counts)
  unsigned long long* s_cand;    // canonical key per slot: original triangle << 32 | candidate
  float4* s_color;               // colour per slot
  uint32_t nfrag;                // slots
  uint32_t nvox;                 // voxels of level 0
};

// ---- exact mean: bricks of the grid accumulated in LDS ----------------------
// The grid is cut into cubes of kBrick^3 voxels (the whole grid when smaller);
// a brick is one contiguous run of the Morton-ordered level 0, and the low
// 3 * log2(kBrick) bits of a voxel's Morton index are its index inside the
// brick.  Triangles are binned to the bricks their voxel bounding box
// overlaps; one workgroup per brick rasterises its triangles' pixels inside
// the brick's footprint on the projection plane, keeps the fragments whose
// voxel lies in the brick, and sums them in LDS as 16.16 fixed point (exact,
// so the order of the additions does not matter).  It then writes the
// brick's level-0 texels (mean = sum / count, 0 where empty) once.  8^3
// bricks with 256-lane workgroups (10 KiB of sums, several workgroups per CU)
// measured 15 % faster than 16^3 bricks with 1024 lanes (80 KiB, one per CU),
// which pay more for workgroup barriers and the load imbalance of the bricks
// than they save on triangles shared by neighbouring bricks.
constexpr int kBrick = 8;
constexpr int kBrickThreads = 256;
// (triangle, brick) items staged per round: a brick holds ~58 items on
// average (1.9 M over 32,768 bricks of a 1M-triangle soup at 256^3), so
// rounds of 128 rarely repeat, and the smaller stage leaves LDS for a 7th
// workgroup per CU (the registers' limit) where 256 allowed 6
constexpr int kBrickItems = 128;

struct BrickGeom {
  uint32_t bits;    // log2(brick edge)
  uint32_t per;     // bricks per grid edge
};

__device__ __forceinline__ BrickGeom brick_geom(uint32_t res) {
  BrickGeom g;
  const uint32_t edge = res < (uint32_t)kBrick ? res : (uint32_t)kBrick;
  g.bits = 31u - __clz(edge);
  g.per = res >> g.bits;
  return g;
}

struct BrickBox {
  int b0[3], nb[3];
  int n;
};

// Bricks overlapped by the voxel bounding box of a set-up triangle.  Fragment
// positions are convex combinations of the vertices up to rounding; 0.01
// voxel of margin covers it.
__device__ __forceinline__ BrickBox brick_box(const float (&pw)[3][3], bool ok, uint32_t res, BrickGeom g) {
#pragma clang fp contract(off)
  BrickBox bb;
  bb.n = 0;
  if (!ok) return bb;
  const float fres = (float)res;
  bb.n = 1;
  for (int k = 0; k < 3; ++k) {
    float lo = 3.0e38f, hi = -3.0e38f;
    for (int v = 0; v < 3; ++v) {
      const float f = fres * (0.5f * pw[v][k] + 0.5f);
      lo = fminf(lo, f);
      hi = fmaxf(hi, f);
    }
    lo -= 0.01f;
    hi += 0.01f;
    if (!(hi > -1.0f && lo < fres)) {  // every fragment would be dropped
      bb.n = 0;
      return bb;
    }
    const int vlo = lo <= 0.0f ? 0 : (int)lo;
    const int vhi = hi >= fres - 1.0f ? (int)res - 1 : (int)hi;
    bb.b0[k] = vlo >> g.bits;
    bb.nb[k] = (vhi >> g.bits) - bb.b0[k] + 1;
    bb.n *= bb.nb[k];
  }
  return bb;
}

__device__ __forceinline__ BrickBox brick_box(const Setup& s, uint32_t res, BrickGeom g) {
  return brick_box(s.w, s.ok, res, g);
}

__device__ __forceinline__ uint32_t brick_of(const BrickBox& bb, int k) {
  const int x = bb.b0[0] + k % bb.nb[0];
  const int y = bb.b0[1] + (k / bb.nb[0]) % bb.nb[1];
  const int z = bb.b0[2] + k / (bb.nb[0] * bb.nb[1]);
  return morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
}

// The brick box of a triangle packed into 8 bytes for the list write: first
// brick per axis, then bricks per axis (0: no brick), 8 bits each (at most
// 128 bricks per axis).
__device__ __forceinline__ uint2 pack_box(const BrickBox& bb) {
  if (bb.n == 0) return make_uint2(0u, 0u);
  return make_uint2((uint32_t)bb.b0[0] | ((uint32_t)bb.b0[1] << 8) | ((uint32_t)bb.b0[2] << 16),
                    (uint32_t)bb.nb[0] | ((uint32_t)bb.nb[1] << 8) | ((uint32_t)bb.nb[2] << 16));
}
__device__ __forceinline__ BrickBox unpack_box(uint2 p) {
  BrickBox bb;
  for (int k = 0; k < 3; ++k) {
    bb.b0[k] = (int)((p.x >> (8 * k)) & 255u);
    bb.nb[k] = (int)((p.y >> (8 * k)) & 255u);
  }
  bb.n = bb.nb[0] * bb.nb[1] * bb.nb[2];
  return bb;
}

// 1. setup (lane per stored triangle): record + count into overlapped bricks.
__global__ __launch_bounds__(256) void vox_brick_count_kernel(DrawTable b, VoxParams vp, TriRec* __restrict__ recs,
                                                              const DevMaterial* __restrict__ mats,
                                                              VoxShade* __restrict__ shade, bool build_shade,
                                                              uint2* __restrict__ boxes, uint32_t* brick_count) {
  // each wave's 64 records (128 B each, consecutive in memory) go out through
  // LDS as lane-contiguous 16-byte stores instead of one record per lane
  // (stride 9: a padded record, so the lanes' 16-byte pieces spread over banks)
  __shared__ uint4 s_out[4][64 * 9];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const BrickGeom g = brick_geom(vp.res);
  Setup s;
  s.ok = false;
  TriRec rec;
  memset(&rec, 0, sizeof(rec));  // a skipped triangle keeps area 0 (set-up triangles have area > 0)
  VoxShade h;
  memset(&h, 0, sizeof(h));
  if (t < b.ntri) {
    const int d = find_draw(b, t);
    const Draw& dr = b.d[d];
    const uint32_t lt = t - dr.tri_base;
    // every stored triangle's material id indexes the table (a default
    // material always exists), so the gather goes out before the set-up's
    // arithmetic instead of after its result
    DevMaterial m;
    if (build_shade) m = mats[dr.mat[lt]];
    tri_setup(dr, lt, vp.res, vp.cube_size, s);
    if (s.ok) {
      rec = make_rec(dr, lt, s);
      if (build_shade) h = make_shade(m);
    }
  }
  {
    uint4* w = s_out[wave];
    const uint4* rp = reinterpret_cast<const uint4*>(&rec);
#pragma unroll
    for (int q = 0; q < 8; ++q) w[lane * 9 + q] = rp[q];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t t0 = t - (uint32_t)lane;
    uint4* out = reinterpret_cast<uint4*>(recs + t0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = k * 64 + lane;  // 16-byte piece c of the wave's 8 KB: record c / 8, piece c % 8
      if (t0 + (uint32_t)(c >> 3) < b.ntri) out[c] = w[(c >> 3) * 9 + (c & 7)];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (build_shade) {
      // the material terms the same way (48 B per lane: a stride of three
      // 16-byte pieces spreads over the banks unpadded)
      const uint4* hp = reinterpret_cast<const uint4*>(&h);
#pragma unroll
      for (int q = 0; q < 3; ++q) w[lane * 3 + q] = hp[q];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      uint4* hout = reinterpret_cast<uint4*>(shade + t0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int c = k * 64 + lane;  // piece c of the wave's 3 KB: triangle c / 3
        if (t0 + (uint32_t)c / 3u < b.ntri) hout[c] = w[c];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
  const BrickBox bb = brick_box(s, vp.res, g);
  if (t < b.ntri) boxes[t] = pack_box(bb);
  for (int k = 0; __ballot(k < bb.n) != 0ull; ++k) {
    const bool act = k < bb.n;
    wave_add_one(brick_count, act ? brick_of(bb, k) : 0u, act);
  }
}

// 3. list write: the same brick boxes, as the setup packed them.  Entries at
//    or past `cap` are not written (their bricks walk the boxes instead).
__global__ __launch_bounds__(256) void vox_brick_write_kernel(uint32_t ntri, const uint2* __restrict__ boxes,
                                                              const uint32_t* __restrict__ brick_off,
                                                              uint32_t* cursor, uint32_t* __restrict__ list,
                                                              uint32_t cap) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const BrickBox bb = unpack_box(t < ntri ? boxes[t] : make_uint2(0u, 0u));
  for (int k = 0; __ballot(k < bb.n) != 0ull; ++k) {
    const bool act = k < bb.n;
    const uint32_t key = act ? brick_of(bb, k) : 0u;
    const uint32_t slot = wave_add_one(cursor, key, act);
    if (act && (uint64_t)brick_off[key] + slot < (uint64_t)cap) list[brick_off[key] + slot] = t;
  }
}

__device__ __forceinline__ uint32_t fixq(float c) { return (uint32_t)(c * 65536.0f + 0.5f); }

// mean of 16.16 fixed-point sums -> stored texel
template <int DT>
__device__ __forceinline__ void store_mean(void* level0, uint32_t v, uint32_t cnt, uint32_t r, uint32_t g, uint32_t b,
                                           uint32_t a) {
#pragma clang fp contract(off)
  float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
  if (cnt) {
    const float fc = (float)cnt;
    const float s = 1.0f / 65536.0f;
    m.x = ((float)r / fc) * s;
    m.y = ((float)g / fc) * s;
    m.z = ((float)b / fc) * s;
    m.w = ((float)a / fc) * s;
  }
  if (DT == VCT_RGBA8) {
    reinterpret_cast<uint32_t*>(level0)[v] =
        to_unorm8(m.x) | (to_unorm8(m.y) << 8) | (to_unorm8(m.z) << 16) | (to_unorm8(m.w) << 24);
  } else if (DT == VCT_FP16) {
    reinterpret_cast<uint2*>(level0)[v] =
        make_uint2((uint32_t)f32_to_f16(m.x) | ((uint32_t)f32_to_f16(m.y) << 16),
                   (uint32_t)f32_to_f16(m.z) | ((uint32_t)f32_to_f16(m.w) << 16));
  } else if (DT == VCT_BF16) {
    reinterpret_cast<uint2*>(level0)[v] =
        make_uint2((uint32_t)f32_to_bf16(m.x) | ((uint32_t)f32_to_bf16(m.y) << 16),
                   (uint32_t)f32_to_bf16(m.z) | ((uint32_t)f32_to_bf16(m.w) << 16));
  } else {
    reinterpret_cast<float4*>(level0)[v] = m;
  }
}

// 4. one workgroup per brick (each workgroup maps its index to one brick's
//    Morton index, one to one).  The
//    brick's triangles are taken 256 at a time; the pixels of their boxes
//    inside the brick's footprint on the projection plane are spread over the
//    lanes (a lane finds its triangle by a binary search over the prefix of
//    the box sizes), coverage is tested from a compact copy of the records in
//    LDS, and only covered pixels read the rest of the record.
struct CovRec {  // the coverage fields of a TriRec
  int64_t area;
  int32_t X[3], Y[3];
  uint32_t flags;
  uint32_t tri;
};

__device__ __forceinline__ void brick_add(uint32_t (*s_sum)[kBrick * kBrick * kBrick], uint32_t* s_cnt, uint32_t v,
                                          float4 color) {
  atomicAdd(&s_sum[0][v], fixq(color.x));
  atomicAdd(&s_sum[1][v], fixq(color.y));
  atomicAdd(&s_sum[2][v], fixq(color.z));
  atomicAdd(&s_sum[3][v], fixq(color.w));
  atomicAdd(&s_cnt[v], 1u);
}

// ---- reference running average -------------------------------------------
__device__ __forceinline__ uint32_t conv_vec4_to_rgba8(float x, float y, float z, float w) {
  // voxelize.frag:66-71: uint() truncation, & 0xFF
  return (((uint32_t)w & 0xffu) << 24) | (((uint32_t)z & 0xffu) << 16) | (((uint32_t)y & 0xffu) << 8) |
         ((uint32_t)x & 0xffu);
}
__device__ __forceinline__ uint32_t enc_nibble(uint32_t m, uint32_t n) {
  return (m & 0xFEFEFEFEu) | (n & 1u) | ((n & 2u) << 7) | ((n & 4u) << 14) | ((n & 8u) << 21);
}
__device__ __forceinline__ uint32_t dec_nibble(uint32_t m) {
  return (m & 1u) | ((m & 0x100u) >> 7) | ((m & 0x10000u) >> 14) | ((m & 0x1000000u) >> 21);
}

// Level-0 texel of one voxel from its fragments in slots [beg, end): sorted
// into canonical order (insertion sort by candidate key), then the packed
// RGBA8 + count-nibble recurrence of imageAtomicRGBA8Avg, one fragment after
// another (voxelize.frag:95-120).
__device__ __forceinline__ uint32_t replay_voxel(unsigned long long* __restrict__ s_cand,
                                                 float4* __restrict__ s_color, uint32_t beg, uint32_t end) {
#pragma clang fp contract(off)
  // canonical order: insertion sort of the bucket by candidate index
  for (uint32_t i = beg + 1; i < end; ++i) {
    const unsigned long long kc = s_cand[i];
    const float4 kv = s_color[i];
    uint32_t j = i;
    while (j > beg && s_cand[j - 1] > kc) {
      s_cand[j] = s_cand[j - 1];
      s_color[j] = s_color[j - 1];
      --j;
    }
    s_cand[j] = kc;
    s_color[j] = kv;
  }
  uint32_t stored = 0u;
  for (uint32_t i = beg; i < end; ++i) {
    const float4 c = s_color[i];
    const float vx = c.x * 255.0f, vy = c.y * 255.0f, vz = c.z * 255.0f, vw = c.w * 255.0f;
    if (stored == 0u) {  // imageAtomicCompSwap(0 -> new) succeeds
      stored = enc_nibble(conv_vec4_to_rgba8(vx, vy, vz, vw), 1u);
      continue;
    }
    const uint32_t cur = stored & 0xFEFEFEFEu;
    float rx = (float)(cur & 0xffu), ry = (float)((cur >> 8) & 0xffu), rz = (float)((cur >> 16) & 0xffu),
          rw = (float)(cur >> 24);
    uint32_t n = dec_nibble(stored);
    const float fn = (float)n;
    rx = rx * fn + vx;
    ry = ry * fn + vy;
    rz = rz * fn + vz;
    rw = rw * fn + vw;
    ++n;
    const float fn1 = (float)n;
    rx = rx / fn1;
    ry = ry / fn1;
    rz = rz / fn1;
    rw = rw / fn1;
    rx = __builtin_rintf(rx / 2.0f) * 2.0f;
    ry = __builtin_rintf(ry / 2.0f) * 2.0f;
    rz = __builtin_rintf(rz / 2.0f) * 2.0f;
    rw = __builtin_rintf(rw / 2.0f) * 2.0f;
    stored = enc_nibble(conv_vec4_to_rgba8(rx, ry, rz, rw), n);
  }
  return stored;
}

// kMode: kExactMean sums the brick's fragments and writes its texels.  The
// reference running average takes two passes over the same candidates:
// kRefCount counts the fragments per voxel (into r.vcount), and after the
// counts are scanned into r.vox_off, kRefPlace shades them and stores
// (canonical key, colour) in the voxel's slots, in no particular order, and
// then replays its brick's voxels and writes their texels.  No global counter
// is shared between workgroups.
// The brick lists' capacity: a brick whose range [off, off + count) runs past
// `cap` entries walks the packed brick boxes of all `ntri` triangles and keeps
// those that contain it (the same triangles its list would hold).
struct ListCap {
  const uint2* boxes;
  uint32_t ntri;
  uint32_t cap;
};

enum { kExactMean = 0, kRefCount = 1, kRefPlace = 2 };
// Its phases wait on LDS and L2 more than they compute (fewer VALU
// instructions per fragment or candidate measured no faster; more waves
// did), so it runs at 8 waves per SIMD: 64 registers, no spills, and LDS for
// 8 workgroups per CU.
template <int DT, int kMode>
__global__ __launch_bounds__(kBrickThreads) __attribute__((amdgpu_waves_per_eu(8))) void vox_brick_kernel(VoxParams vp, const TriRec* __restrict__ recs,
                                                                  const VoxShade* __restrict__ shade,
                                                                  const uint32_t* __restrict__ brick_count,
                                                                  const uint32_t* __restrict__ brick_off,
                                                                  const uint32_t* __restrict__ list,
                                                                  ListCap lc, void* level0, RefPass r) {
  constexpr int kVox = kBrick * kBrick * kBrick;
  __shared__ uint32_t s_sum[4][kMode == kExactMean ? kVox : 1];
  __shared__ uint32_t s_cnt[kVox];
  __shared__ uint32_t s_base[kMode == kRefPlace ? kVox : 1];
  __shared__ CovRec s_cov[kBrickItems];
  __shared__ uint32_t s_end[kBrickItems];  // inclusive prefix of candidate counts
  __shared__ uint32_t s_org[kBrickItems];  // clipped box origin: i | j << 16
  __shared__ uint32_t s_div[kBrickItems];  // ceil(65536 / width) | width << 20
  __shared__ uint32_t s_wsum[kBrickItems / 64];
  __shared__ uint32_t s_q[kBrickThreads / 64][128];  // per-wave queue of covered candidates
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const BrickGeom g = brick_geom(vp.res);
  // Workgroups are dealt to the 8 XCDs round-robin in dispatch order; XCD k
  // takes the k-th eighth of the bricks in Morton order (one region of the
  // grid), so the records of triangles shared by neighbouring bricks stay
  // in one L2 (voxelize 1.04 -> 1.01 ms)
  uint32_t brick = blockIdx.x;
  {
    const uint32_t full = gridDim.x & ~7u;
    if (brick < full) brick = (brick & 7u) * (full >> 3) + (brick >> 3);
  }
  const uint32_t nvox = 1u << (3 * g.bits);
  const uint32_t bshift = 3 * g.bits;
  const uint32_t bc[3] = {compact3(brick), compact3(brick >> 1), compact3(brick >> 2)};
  for (uint32_t v = tid; v < nvox; v += kBrickThreads) {
    if constexpr (kMode == kExactMean) {
      s_sum[0][v] = 0u;
      s_sum[1][v] = 0u;
      s_sum[2][v] = 0u;
      s_sum[3][v] = 0u;
    }
    if constexpr (kMode == kRefPlace)
      s_base[v] = r.vox_off[brick * nvox + v];
    s_cnt[v] = 0u;
  }
  const uint32_t n = brick_count[brick], base = brick_off[brick];
  // a brick whose list runs past the lists' capacity (the geometry grew past
  // what the lists were sized for) walks every triangle's brick box instead
  const bool walk = (uint64_t)base + n > (uint64_t)lc.cap;
  const uint32_t nitems = walk ? lc.ntri : n;
  const int lim = 2 * (int)vp.res - 1;
  const int span = 2 << g.bits;  // footprint of the brick on the raster, pixels per axis
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nitems; b0 += kBrickItems) {
    uint32_t cnt = 0;
    uint32_t t = b0 + (uint32_t)tid;
    bool in = tid < kBrickItems && t < nitems;
    if (in && walk) {  // the box as pack_box stores it: first brick, bricks per axis
      const uint2 pb = lc.boxes[t];
      for (int k = 0; k < 3; ++k)
        in = in && bc[k] - ((pb.x >> (8 * k)) & 255u) < ((pb.y >> (8 * k)) & 255u);
    } else if (in) {
      t = list[base + t];
    }
    if (in) {
      const TriRec* rp = recs + t;
      const uint32_t fl = rp->flags;
      const int axis = (int)((fl >> 1) & 3u);
      const int ia = axis == 0 ? 1 : 0, ib = axis == 2 ? 1 : 2;
      const int qh = (int)(fl >> 8);
      const int qx0 = rp->qx0, qy0 = rp->qy0, qw = rp->qw;
      // the brick's footprint clipped to the triangle's box.  A covered
      // pixel (i, j) lands in voxel (i >> 1, j >> 1) on the projection plane:
      // its centre sits a quarter voxel inside that voxel, and the
      // interpolated position is a convex combination of the vertices that
      // differs from the centre by the 1/512-pixel vertex snap plus rounding,
      // far below a quarter voxel.  So pixels outside the footprint never
      // land in the brick.
      const int fx = (int)bc[ia] * span, fy = (int)bc[ib] * span;
      const int i0 = max(qx0, fx), i1 = min(min(qx0 + 2 * qw - 1, fx + span - 1), lim);
      const int j0 = max(qy0, fy), j1 = min(min(qy0 + 2 * qh - 1, fy + span - 1), lim);
      if (i1 >= i0 && j1 >= j0) {
        const uint32_t cw = (uint32_t)(i1 - i0 + 1);
        cnt = cw * (uint32_t)(j1 - j0 + 1);
        s_org[tid] = (uint32_t)i0 | ((uint32_t)j0 << 16);
        s_div[tid] = ((65536u + cw - 1u) / cw) | (cw << 20);
        CovRec cr;
        cr.area = rp->area;
        for (int v = 0; v < 3; ++v) {
          cr.X[v] = rp->X[v];
          cr.Y[v] = rp->Y[v];
        }
        cr.flags = fl;
        cr.tri = t;
        s_cov[tid] = cr;
      }
    }
    uint32_t x = cnt;
    if (tid < kBrickItems) {
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane == 63) s_wsum[wave] = x;
    }
    __syncthreads();
    if (tid < kBrickItems) {
      for (int w = 0; w < wave; ++w) x += s_wsum[w];
      s_end[tid] = x;
    }
    __syncthreads();
    const uint32_t total = s_end[kBrickItems - 1];
    // Candidates in two phases.  1: the wave's 64 candidates are tested for
    // coverage only (exact edge functions, top-left rule); the covered ones
    // are appended to the wave's queue in LDS.  2: whenever the queue holds
    // 64, every lane takes one and runs the exact fragment path (voxel
    // address, brick test, shading, accumulation), so the expensive part
    // runs with all lanes busy instead of on the covered ~40 % of them.  The
    // set of fragments, and every value computed for one, is the same as
    // testing each candidate in full.
    uint32_t* q = s_q[wave];
    uint32_t qn = 0;  // entries in the queue (the same in every lane)
    uint32_t own0 = 0;  // the entry of the wave's first candidate (the same in every lane; never decreases)
    auto fragment = [&](uint32_t d) {
      const uint32_t lo = d & 255u;
      const int i = (int)((d >> 8) & 0x7ffu), j = (int)(d >> 19);
      const CovRec cr = s_cov[lo];
      VEdges e;
      vox_edges(cr, i, j, e);
      uint32_t key;
      float4 color;
      if (!shade_w<true, kMode == kRefCount>(cr, recs + cr.tri, shade + cr.tri, e, i, j, e.w[0], e.w[1], e.w[2], vp, key,
                                             color, bshift, brick))
        return;
      const uint32_t lv = key & (nvox - 1u);
      if constexpr (kMode == kExactMean) {
        brick_add(s_sum, s_cnt, lv, color);
      } else if constexpr (kMode == kRefCount) {
        atomicAdd(&s_cnt[lv], 1u);
      } else {
        // canonical key: the candidate's place in the triangle's quad-ordered
        // box (2x2 quads row by row, each quad x then y)
        const TriRec* rp = recs + cr.tri;
        const uint32_t dx = (uint32_t)(i - rp->qx0), dy = (uint32_t)(j - rp->qy0);
        const uint32_t lc = (((dy >> 1) * (uint32_t)rp->qw + (dx >> 1)) << 2) | ((dy & 1u) << 1) | (dx & 1u);
        const uint32_t slot = s_base[lv] + atomicAdd(&s_cnt[lv], 1u);
        if (slot < r.nfrag) {  // always: the count pass found the same fragments
          r.s_cand[slot] = ((unsigned long long)rp->otri << 32) | (unsigned long long)lc;
          r.s_color[slot] = color;
        }
      }
    };
    for (uint32_t cb = (uint32_t)wave * 64u; cb < total; cb += kBrickThreads) {
      const uint32_t c = cb + (uint32_t)lane;
      // The entry (triangle) of each lane's candidate, for the window
      // [cb, cb + 64) of consecutive candidates.  own0 moves to the first
      // entry ending after cb: lane k tests entry own0 + k and the ballot
      // counts those ending at or before cb (a prefix, the ends being
      // sorted).  Lane k then holds the end of entry own0 + k, and each lane
      // finds how many of these end at or before its candidate by a binary
      // search over the lanes (register shuffles): that count from own0 is
      // its entry.  A window that spans more than 64 entries (entries with
      // one candidate or none) finishes with a per-lane walk.
      for (;;) {
        const uint32_t ek = own0 + (uint32_t)lane < kBrickItems ? s_end[own0 + lane] : 0xffffffffu;
        const uint32_t k = (uint32_t)__popcll(__ballot(ek <= cb));
        own0 += k;
        if (k < 64u) break;
      }
      const uint32_t ek = own0 + (uint32_t)lane < kBrickItems ? s_end[own0 + lane] : 0xffffffffu;
      // Only entries ending inside the window can end at or before a
      // candidate: nend of them, the first nend lanes (the ends are sorted).  A
      // window usually spans a few entries, and a scan of their ends read
      // into scalar registers has no LDS round trip in its chain, where the
      // binary search takes seven dependent shuffles; the count is the same.
      const uint32_t nend = (uint32_t)__popcll(__ballot(ek < cb + 64u));
      uint32_t past = 0;
      if (nend <= 8u) {
        for (uint32_t k = 0; k < nend; ++k) past += (uint32_t)__builtin_amdgcn_readlane((int)ek, (int)k) <= c ? 1u : 0u;
      } else {
        for (uint32_t st = 32; st; st >>= 1)
          if ((uint32_t)__shfl(ek, (int)(past + st - 1u), 64) <= c) past += st;
        if ((uint32_t)__shfl(ek, (int)past, 64) <= c) ++past;
      }
      bool hit = false;
      uint32_t desc = 0;
      if (c < total) {
        uint32_t lo = own0 + past;
        if (past == 64u)
          while (s_end[lo] <= c) ++lo;
        const uint32_t local = c - (lo ? s_end[lo - 1] : 0u);
        const uint32_t dv = s_div[lo], cw = dv >> 20;
        const uint32_t row = (local * (dv & 0xfffffu)) >> 16;  // exact: local < 35 * 35, cw <= 35
        const uint32_t col = local - row * cw;
        const uint32_t org = s_org[lo];
        const int i = (int)(org & 0xffffu) + (int)col, j = (int)(org >> 16) + (int)row;
        const CovRec cr = s_cov[lo];
        VEdges e;
        vox_edges(cr, i, j, e);
        hit = vox_covers(e);
        desc = lo | ((uint32_t)i << 8) | ((uint32_t)j << 19);
      }
      const unsigned long long m = __ballot(hit);
      if (hit) q[qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = desc;
      qn += (uint32_t)__popcll(m);
      if (qn >= 64u) {
        qn -= 64u;
        fragment(q[qn + (uint32_t)lane]);
      }
    }
    if ((uint32_t)lane < qn) fragment(q[lane]);
    __syncthreads();
  }
  if constexpr (kMode == kExactMean) {
    for (uint32_t v = tid; v < nvox; v += kBrickThreads)
      store_mean<DT>(level0, brick * nvox + v, s_cnt[v], s_sum[0][v], s_sum[1][v], s_sum[2][v], s_sum[3][v]);
  } else if (kMode == kRefCount) {
    // one workgroup per brick: it writes the count of every voxel it owns
    for (uint32_t v = tid; v < nvox; v += kBrickThreads) r.vcount[brick * nvox + v] = s_cnt[v];
  } else {
    // the barrier makes the workgroup's slot stores visible to all its lanes
    // (the loop above ended with one)
    for (uint32_t v = tid; v < nvox; v += kBrickThreads) {
      const uint32_t gv = brick * nvox + v;
      const uint32_t beg = r.vox_off[gv];
      const uint32_t end = gv + 1u < r.nvox ? r.vox_off[gv + 1u] : r.nfrag;
      reinterpret_cast<uint32_t*>(level0)[gv] = replay_voxel(r.s_cand, r.s_color, beg, end);
    }
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace vct

using namespace vct;

// mat3(transpose(inverse(M))) of a column-major 4x4 (voxelize.vert:28).
static void normal_matrix(const float* M, float* N) {
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
  // transpose(inv), column-major: N[c*3 + r] = inv^T[r][c] = inv[c][r]
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) N[c * 3 + r] = (float)inv[c][r];
}

static LightSet make_lights(const vct_context* ctx) {
  LightSet ls;
  memset(&ls, 0, sizeof(ls));
  int n = (int)ctx->lights.size();
  if (n > VCT_MAX_LIGHTS) n = VCT_MAX_LIGHTS;
  for (int i = 0; i < n; ++i) {
    const vct_light& L = ctx->lights[i];
    for (int k = 0; k < 3; ++k) {
      ls.l[i].pos[k] = L.position[k];
      ls.l[i].color[k] = L.color[k];
    }
    ls.l[i].intensity = L.intensity;
    ls.l[i].kind = L.kind;
  }
  ls.count = n;
  return ls;
}

// Every queued draw into the device draw table (through pinned staging, in
// stream order with the kernels that read the previous frame's table).
static int fill_draws(vct_context* ctx, const std::vector<uint32_t>& q, DrawTable& b, bool upload = true) {
  memset(&b, 0, sizeof(b));
  if (q.size() > (size_t)0x7fffffff) return vct_set_error(VCT_ERR_LIMIT, "too many queued models");
  std::vector<Draw> draws(q.size());
  uint64_t ntri = 0;
  uint32_t canon = 0;
  for (size_t k = 0; k < q.size(); ++k) {
    const VctModel& m = ctx->models[q[k]];
    Draw& d = draws[k];
    d.pos = m.d_pos;
    d.nrm = m.d_nrm;
    d.mat = m.d_mat;
    d.perm = m.d_perm;
    d.tri_base = ntri;
    d.canon_base = canon;
    canon += m.ntri;
    d.ntri = m.ntri;
    memcpy(d.M, m.xform, 64);
    normal_matrix(m.xform, d.N);
    ntri += m.ntri;
  }
  if (ntri > (1ull << 31)) return vct_set_error(VCT_ERR_LIMIT, "more than 2^31 triangles queued");
  b.ndraw = (int)draws.size();
  b.ntri = (uint32_t)ntri;
  int rc;
  if (upload && b.ndraw && (rc = vct_upload(ctx, ctx->v_draws, draws.data(), draws.size() * sizeof(Draw))))
    return rc;
  b.d = (const Draw*)ctx->v_draws.dev.p;
  return VCT_OK;
}

static size_t setup_bytes(uint32_t ntri) { return (size_t)(ntri ? ntri : 1) * sizeof(TriRec); }

// The set-up launch; it also writes the material terms of the frame's
// triangles unless those of the last build still hold: the same queued
// models (so the same frame order) at the same material epoch, in the same
// buffer.
static int launch_setup(vct_context* ctx, const std::vector<uint32_t>& q, const DrawTable& b, const VoxParams& vp,
                        uint32_t* bcount) {
  int rc;
  if ((rc = vct_ensure(ctx->vox_shade, (size_t)(b.ntri ? b.ntri : 1) * sizeof(VoxShade), ctx))) return rc;
  const bool build = !(ctx->vox_shade_epoch == ctx->shade_epoch && ctx->vox_shade_models == q &&
                       ctx->vox_shade_ptr == ctx->vox_shade.p);
  hipLaunchKernelGGL(vox_brick_count_kernel, dim3((b.ntri + 255) / 256), dim3(256), 0, ctx->stream, b, vp,
                     (TriRec*)ctx->setup.p, (const DevMaterial*)ctx->d_materials.p, (VoxShade*)ctx->vox_shade.p,
                     build, (uint2*)ctx->vox_box.p, bcount);
  VCT_HIP(hipGetLastError());
  ctx->vox_shade_epoch = ctx->shade_epoch;
  ctx->vox_shade_models = q;
  ctx->vox_shade_ptr = ctx->vox_shade.p;
  return VCT_OK;
}

template <int DT, int kMode = kExactMean>
static void launch_brick(vct_context* ctx, uint32_t nbricks, uint32_t ntri, const VoxParams& vp,
                         const uint32_t* count, const uint32_t* off, const uint32_t* list, void* level0,
                         const RefPass& r = RefPass{}, ListCap lc = ListCap{nullptr, 0u, 0xffffffffu}) {
  hipLaunchKernelGGL((vox_brick_kernel<DT, kMode>), dim3(nbricks), dim3(kBrickThreads), 0, ctx->stream, vp,
                     (const TriRec*)ctx->setup.p, (const VoxShade*)ctx->vox_shade.p, count, off, list, lc, level0,
                     r);
}

// Exact mean: all queued draws in one batch (brick lists span every draw).
// Once a (triangle, brick) pair count is known for this many triangles, phase
// 1 launches everything: the list is sized from that count with a margin
// (a brick whose range runs past it walks the triangles' boxes, so a frame
// whose geometry grew is still exact), and this frame's count comes back
// asynchronously for the next.  Otherwise (the first frame, or another
// triangle count) phase 1 launches up to the copy of the count to pinned
// host memory and phase 2 waits for it and launches the rest.
static int voxelize_exact(vct_context* ctx, const std::vector<uint32_t>& q, const VoxParams& vp, void* level0,
                          int phase) {
  const uint32_t res = ctx->res;
  const uint32_t edge = res < (uint32_t)kBrick ? res : (uint32_t)kBrick;
  const uint32_t per = res / edge;
  const uint32_t nbricks = per * per * per;
  uint32_t* counters = (uint32_t*)ctx->counters.p;
  int rc;
  // [0, nbricks): counts, [nbricks, 2 nbricks): cursors; then offsets
  if ((rc = vct_ensure(ctx->vox_count, (size_t)nbricks * 8 + 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->vox_off, (size_t)nbricks * 4 + 16, ctx))) return rc;
  uint32_t* bcount = (uint32_t*)ctx->vox_count.p;
  uint32_t* cursor = bcount + nbricks;
  uint32_t* boff = (uint32_t*)ctx->vox_off.p;
  DrawTable b;
  if ((rc = fill_draws(ctx, q, b, (phase & 1) != 0))) return rc;
  const uint32_t ntri = b.ntri;
  auto bricks = [&](const uint32_t* list, ListCap lc) {
    switch (ctx->dtype) {
      case VCT_RGBA8: launch_brick<VCT_RGBA8>(ctx, nbricks, ntri, vp, bcount, boff, list, level0, RefPass{}, lc); break;
      case VCT_FP16: launch_brick<VCT_FP16>(ctx, nbricks, ntri, vp, bcount, boff, list, level0, RefPass{}, lc); break;
      case VCT_BF16: launch_brick<VCT_BF16>(ctx, nbricks, ntri, vp, bcount, boff, list, level0, RefPass{}, lc); break;
      default: launch_brick<VCT_FP32>(ctx, nbricks, ntri, vp, bcount, boff, list, level0, RefPass{}, lc); break;
    }
  };
  if (phase & 1) {
    // the count an earlier frame copied back, if it has landed (never waited for)
    if (ctx->vox_total_pending) {
      const hipError_t qe = hipEventQuery(ctx->vox_total_ev);
      if (qe == hipSuccess) {
        ctx->vox_total_pending = false;
        ctx->vox_known_total = ctx->h_readback[3];
      } else if (qe == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: the copy is still in flight
      } else {
        return vct_hip_check(qe, "hipEventQuery");
      }
    }
    hipLaunchKernelGGL(zero_u32_kernel, dim3((2 * nbricks + 255) / 256), dim3(256), 0, ctx->stream, bcount,
                       (uint64_t)2 * nbricks);
    VCT_HIP(hipGetLastError());
    ctx->h_readback[2] = 0u;
    ctx->vox_sized = ntri > 0 && ctx->vox_known_total > 0 && ctx->vox_known_ntri == ntri;
    if (ntri) {
      if ((rc = vct_ensure(ctx->setup, setup_bytes(ntri), ctx))) return rc;
      if ((rc = vct_ensure(ctx->vox_box, (size_t)ntri * 8, ctx))) return rc;
      if ((rc = launch_setup(ctx, q, b, vp, bcount))) return rc;
      if ((rc = vct_launch_scan_u32(ctx, bcount, boff, nbricks, counters + 2))) return rc;
      if (!ctx->vox_sized) {
        VCT_HIP(hipMemcpyAsync(ctx->h_readback + 2, counters + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
      } else {
        // a quarter more than the last count, at least two pairs per triangle
        const uint64_t want = std::max<uint64_t>((uint64_t)ctx->vox_known_total + ctx->vox_known_total / 4,
                                                 2ull * ntri + 1024u);
        const uint32_t cap = (uint32_t)std::min<uint64_t>(want, 0xffffffffull);
        if ((rc = vct_ensure(ctx->tri_off, (size_t)cap * 4, ctx))) return rc;
        ctx->vox_cap = cap;
        hipLaunchKernelGGL(vox_brick_write_kernel, dim3((ntri + 255) / 256), dim3(256), 0, ctx->stream, ntri,
                           (const uint2*)ctx->vox_box.p, (const uint32_t*)boff, cursor, (uint32_t*)ctx->tri_off.p,
                           cap);
        VCT_HIP(hipGetLastError());
        VCT_HIP(hipMemcpyAsync(ctx->h_readback + 3, counters + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
        VCT_HIP(hipEventRecord(ctx->vox_total_ev, ctx->stream));
        ctx->vox_total_pending = true;
        ctx->vox_pairs_on_device = true;
        bricks((const uint32_t*)ctx->tri_off.p, ListCap{(const uint2*)ctx->vox_box.p, ntri, cap});
        VCT_HIP(hipGetLastError());
      }
    }
  }
  if (!(phase & 2) || ctx->vox_sized) return VCT_OK;
  // also when nothing was counted: a copy an interrupted earlier frame left
  // in flight must not land after the word is read
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t total = ctx->h_readback[2];
  if (total) {
    if ((rc = vct_ensure(ctx->tri_off, (size_t)total * 4 + (size_t)total / 4 * 4, ctx))) return rc;
    hipLaunchKernelGGL(vox_brick_write_kernel, dim3((ntri + 255) / 256), dim3(256), 0, ctx->stream, ntri,
                       (const uint2*)ctx->vox_box.p, (const uint32_t*)boff, cursor, (uint32_t*)ctx->tri_off.p,
                       total);
    VCT_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL(zero_u32_kernel, dim3((nbricks + 255) / 256), dim3(256), 0, ctx->stream, boff,
                       (uint64_t)nbricks);
    VCT_HIP(hipGetLastError());
    if ((rc = vct_ensure(ctx->tri_off, 16, ctx))) return rc;
  }
  if ((rc = vct_ensure(ctx->setup, setup_bytes(ntri), ctx))) return rc;
  ctx->last_candidates = total;  // (triangle, brick) pairs
  ctx->vox_known_total = total;
  ctx->vox_known_ntri = ntri;
  bricks((const uint32_t*)ctx->tri_off.p, ListCap{nullptr, 0u, 0xffffffffu});
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// Brick lists of the queued draws for the reference running average:
// records, per-brick triangle counts (bcount), offsets (boff) and lists
// (tri_off).  Returns the number of (triangle, brick) pairs in `pairs`.
static int ref_lists(vct_context* ctx, const std::vector<uint32_t>& q, const DrawTable& b, const VoxParams& vp,
                     uint32_t nbricks, uint32_t& pairs) {
  uint32_t* counters = (uint32_t*)ctx->counters.p;
  uint32_t* bcount = (uint32_t*)ctx->tri_count.p;
  uint32_t* cursor = bcount + nbricks;
  uint32_t* boff = (uint32_t*)ctx->tri_base.p;
  const uint32_t ntri = b.ntri;
  int rc;
  pairs = 0;
  if ((rc = vct_ensure(ctx->setup, setup_bytes(ntri), ctx))) return rc;
  if ((rc = vct_ensure(ctx->tri_off, 16, ctx))) return rc;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((2 * nbricks + 255) / 256), dim3(256), 0, ctx->stream, bcount,
                     (uint64_t)2 * nbricks);
  VCT_HIP(hipGetLastError());
  if (!ntri) return VCT_OK;
  if ((rc = vct_ensure(ctx->vox_box, (size_t)ntri * 8, ctx))) return rc;
  if ((rc = launch_setup(ctx, q, b, vp, bcount))) return rc;
  if ((rc = vct_launch_scan_u32(ctx, bcount, boff, nbricks, counters + 2))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->h_readback + 2, counters + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  pairs = ctx->h_readback[2];
  if (!pairs) return VCT_OK;
  if ((rc = vct_ensure(ctx->tri_off, (size_t)pairs * 4 + (size_t)pairs / 4 * 4, ctx))) return rc;
  hipLaunchKernelGGL(vox_brick_write_kernel, dim3((ntri + 255) / 256), dim3(256), 0, ctx->stream, ntri,
                     (const uint2*)ctx->vox_box.p, (const uint32_t*)boff, cursor, (uint32_t*)ctx->tri_off.p, pairs);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// Reference running average: the brick pipeline of the exact mean finds the
// same fragments (each covered pixel of a triangle once, in the brick of its
// voxel).  A count pass writes the fragments of every voxel, a scan gives
// every voxel its slots, a place pass stores (canonical key, colour) in them,
// and the replay sorts each voxel's slots and runs the RGBA8 recurrence.
static int voxelize_reference(vct_context* ctx, const std::vector<uint32_t>& q, const VoxParams& vp, void* level0) {
  const uint32_t res = ctx->res;
  const uint64_t nvox = (uint64_t)res * res * res;
  const uint32_t edge = res < (uint32_t)kBrick ? res : (uint32_t)kBrick;
  const uint32_t per = res / edge;
  const uint32_t nbricks = per * per * per;
  uint32_t* counters = (uint32_t*)ctx->counters.p;  // [1] fragments, [2] scan totals
  int rc;
  if ((rc = vct_ensure(ctx->vox_count, nvox * 4, ctx))) return rc;
  if ((rc = vct_ensure(ctx->vox_off, nvox * 4, ctx))) return rc;
  // per-brick counts + cursors, per-brick offsets
  if ((rc = vct_ensure(ctx->tri_count, (size_t)nbricks * 8 + 16, ctx))) return rc;
  if ((rc = vct_ensure(ctx->tri_base, (size_t)nbricks * 4 + 16, ctx))) return rc;
  uint32_t* vcount = (uint32_t*)ctx->vox_count.p;
  const uint32_t* bcount = (const uint32_t*)ctx->tri_count.p;
  const uint32_t* boff = (const uint32_t*)ctx->tri_base.p;
  DrawTable b;
  if ((rc = fill_draws(ctx, q, b))) return rc;
  uint32_t pairs = 0;
  if ((rc = ref_lists(ctx, q, b, vp, nbricks, pairs))) return rc;
  ctx->last_candidates += pairs;
  RefPass r;
  memset(&r, 0, sizeof(r));
  r.vcount = vcount;
  // every brick's workgroup writes its voxels' counts, zeros included
  launch_brick<VCT_RGBA8, kRefCount>(ctx, nbricks, b.ntri, vp, bcount, boff, (const uint32_t*)ctx->tri_off.p, nullptr,
                                     r);
  VCT_HIP(hipGetLastError());
  if ((rc = vct_launch_scan_u32(ctx, vcount, (uint32_t*)ctx->vox_off.p, (uint32_t)nvox, counters + 1))) return rc;
  VCT_HIP(hipMemcpyAsync(ctx->h_readback + 2, counters + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  VCT_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t nfrag = ctx->h_readback[2];
  ctx->last_fragments = nfrag;
  if (!nfrag) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(2048), dim3(256), 0, ctx->stream, (uint32_t*)level0, nvox);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
  }
  const size_t ns = ((size_t)nfrag + 4) & ~(size_t)3;
  if ((rc = vct_ensure(ctx->frags_sorted, ns * 24, ctx))) return rc;
  r.vox_off = (const uint32_t*)ctx->vox_off.p;
  r.s_cand = (unsigned long long*)ctx->frags_sorted.p;
  r.s_color = (float4*)((char*)ctx->frags_sorted.p + ns * 8);
  r.nfrag = nfrag;
  r.nvox = (uint32_t)nvox;
  launch_brick<VCT_RGBA8, kRefPlace>(ctx, nbricks, b.ntri, vp, bcount, boff, (const uint32_t*)ctx->tri_off.p, level0,
                                     r);
  VCT_HIP(hipGetLastError());
  return VCT_OK;
}

// normalize(position) of the directional lights, operation for operation as
// the fragment shader computes it (voxelize.frag:127-131), without contraction.
static void light_dirs(VoxParams& vp) {
#pragma clang fp contract(off)
  for (int li = 0; li < VCT_MAX_LIGHTS; ++li) {
    const DevLight& L = vp.lights.l[li];
    const float l = sqrtf(L.pos[0] * L.pos[0] + L.pos[1] * L.pos[1] + L.pos[2] * L.pos[2]);
    for (int k = 0; k < 3; ++k) vp.ldir[li][k] = L.pos[k] / l;
  }
}

int vct_launch_voxelize(vct_context* ctx) { return vct_launch_voxelize_phase(ctx, 3); }

int vct_launch_voxelize_phase(vct_context* ctx, int phase) {
  const uint32_t res = ctx->res;
  void* level0 = (char*)ctx->volume.p + ctx->lvl_off[0][0];
  int rc;
  if ((rc = vct_ensure(ctx->counters, 256, ctx))) return rc;
  VoxParams vp;
  memset(&vp, 0, sizeof(vp));
  vp.res = res;
  vp.cube_size = ctx->cube_size;
  vp.average = ctx->average;
  vp.lights = make_lights(ctx);
  light_dirs(vp);
  std::vector<uint32_t> q;
  for (uint32_t id : ctx->queue)
    if (id < ctx->models.size() && ctx->models[id].ntri) q.push_back(id);
  if (phase & 1) {
    ctx->last_candidates = 0;
    ctx->last_fragments = 0;
    ctx->vox_pairs_on_device = false;
  }
  if (ctx->average == VCT_AVG_EXACT) return voxelize_exact(ctx, q, vp, level0, phase);
  return (phase & 2) ? voxelize_reference(ctx, q, vp, level0) : VCT_OK;
}
