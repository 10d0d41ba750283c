// vct_scene.cpp - deterministic procedural scenes (host code).
//
// The reference renders the Cornell-Box-Glossy + Suzanne OBJ scene set up in
// main.cpp:85-121.  The workloads measured here are procedural so they can be
// generated at any size without files: a Cornell box of 34 triangles and
// random triangle soups.  Random numbers: SplitMix64 seeds a PCG32 (XSH RR)
// stream; every derived quantity is computed in double, then stored as float.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/vct.h"

namespace {

struct Pcg32 {
  uint64_t state, inc;
  explicit Pcg32(uint64_t seed) {
    uint64_t z = seed;
    auto splitmix = [&z]() {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t r = z;
      r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
      r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
      return r ^ (r >> 31);
    };
    state = splitmix();
    inc = splitmix() | 1ull;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  // uniform in [0, 1) with 53 bits
  double uniform() {
    uint64_t hi = next(), lo = next();
    return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
  }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

void rotate(const double q[4], const double v[3], double out[3]) {
  // q = (x, y, z, w), unit
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double r[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                          {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                          {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
  for (int i = 0; i < 3; ++i) out[i] = r[i][0] * v[0] + r[i][1] * v[1] + r[i][2] * v[2];
}

}  // namespace

extern "C" {

int vct_scene_soup(uint64_t seed, uint32_t ntri, float extent, float mu, float sigma, float rmin, float rmax,
                   float* pos, float* nrm, float* albedo_rough) {
  if ((ntri && (!pos || !nrm)) || !(extent > 0.0f) || !(sigma >= 0.0f) || rmin > rmax) return VCT_ERR_INVALID_ARG;
  Pcg32 rng(seed);
  const double r3 = std::sqrt(3.0);
  for (uint32_t t = 0; t < ntri; ++t) {
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = (2.0 * rng.uniform() - 1.0) * extent;
    const double edge = std::exp((double)mu + (double)sigma * rng.normal());
    // uniform rotation (Shoemake)
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    const double q[4] = {a * std::sin(6.283185307179586 * u2), a * std::cos(6.283185307179586 * u2),
                         b * std::sin(6.283185307179586 * u3), b * std::cos(6.283185307179586 * u3)};
    const double R = edge / r3;  // circumradius of the equilateral triangle
    const double ang[3] = {1.5707963267948966, 1.5707963267948966 + 2.0943951023931953,
                           1.5707963267948966 + 4.1887902047863905};
    double n[3];
    const double z[3] = {0.0, 0.0, 1.0};
    rotate(q, z, n);
    for (int v = 0; v < 3; ++v) {
      const double l[3] = {R * std::cos(ang[v]), R * std::sin(ang[v]), 0.0};
      double w[3];
      rotate(q, l, w);
      for (int k = 0; k < 3; ++k) {
        pos[(size_t)t * 9 + v * 3 + k] = (float)(c[k] + w[k]);
        nrm[(size_t)t * 9 + v * 3 + k] = (float)n[k];
      }
    }
    const double alb[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
    const double rough = rmin + (rmax - rmin) * rng.uniform();
    if (albedo_rough) {
      for (int k = 0; k < 3; ++k) albedo_rough[(size_t)t * 4 + k] = (float)alb[k];
      albedo_rough[(size_t)t * 4 + 3] = (float)rough;
    }
  }
  return VCT_OK;
}

int vct_scene_cornell(uint64_t seed, float* pos, float* nrm, uint32_t* tri_material, vct_material* mats) {
  if (!pos || !nrm || !tri_material || !mats) return VCT_ERR_INVALID_ARG;
  Pcg32 rng(seed);
  int t = 0;
  auto tri = [&](const double* a, const double* b, const double* c, const double* n, uint32_t m) {
    const double* v[3] = {a, b, c};
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        pos[t * 9 + i * 3 + k] = (float)v[i][k];
        nrm[t * 9 + i * 3 + k] = (float)n[k];
      }
    tri_material[t] = m;
    ++t;
  };
  auto quad = [&](const double* a, const double* b, const double* c, const double* d, const double* n,
                  uint32_t m) {
    tri(a, b, c, n, m);
    tri(a, c, d, n, m);
  };
  const double h = 0.5;
  // room corners
  const double p000[3] = {-h, -h, -h}, p100[3] = {h, -h, -h}, p010[3] = {-h, h, -h}, p110[3] = {h, h, -h};
  const double p001[3] = {-h, -h, h}, p101[3] = {h, -h, h}, p011[3] = {-h, h, h}, p111[3] = {h, h, h};
  const double ny[3] = {0, 1, 0}, nyn[3] = {0, -1, 0}, nz[3] = {0, 0, 1}, nx[3] = {1, 0, 0}, nxn[3] = {-1, 0, 0};
  quad(p000, p001, p101, p100, ny, 0);   // floor (white), normal +y
  quad(p010, p110, p111, p011, nyn, 0);  // ceiling (white), normal -y
  quad(p000, p100, p110, p010, nz, 0);   // back wall z = -h (white), normal +z
  quad(p000, p010, p011, p001, nx, 1);   // left wall x = -h (red), normal +x
  quad(p100, p101, p111, p110, nxn, 2);  // right wall x = +h (green), normal -x
  // two boxes rotated about y, resting on the floor
  struct Box {
    double cx, cz, sx, sy, sz, ang;
    uint32_t m;
  } boxes[2] = {{-0.16, -0.14, 0.28, 0.58, 0.28, 0.30, 3}, {0.17, 0.14, 0.28, 0.28, 0.28, -0.29, 4}};
  for (const Box& b : boxes) {
    const double c = std::cos(b.ang), s = std::sin(b.ang);
    auto P = [&](double x, double y, double z, double* out) {
      // local (x,z) rotated about y, y from the floor up
      out[0] = b.cx + c * x + s * z;
      out[1] = -h + y;
      out[2] = b.cz - s * x + c * z;
    };
    double v[8][3];
    for (int i = 0; i < 8; ++i)
      P(((i & 1) ? 0.5 : -0.5) * b.sx, ((i & 2) ? 1.0 : 0.0) * b.sy, ((i & 4) ? 0.5 : -0.5) * b.sz, v[i]);
    const double ex[3] = {c, 0, -s}, ez[3] = {s, 0, c};
    const double mex[3] = {-c, 0, s}, mez[3] = {-s, 0, -c};
    quad(v[0], v[1], v[3], v[2], mez, b.m);  // z- face
    quad(v[4], v[6], v[7], v[5], ez, b.m);   // z+ face
    quad(v[0], v[2], v[6], v[4], mex, b.m);  // x- face
    quad(v[1], v[5], v[7], v[3], ex, b.m);   // x+ face
    quad(v[2], v[3], v[7], v[6], ny, b.m);   // top
    quad(v[0], v[4], v[5], v[1], nyn, b.m);  // bottom
  }
  // materials: 0 white, 1 red, 2 green, 3-4 boxes with seeded albedo
  const double col[5][3] = {{0.73, 0.73, 0.73}, {0.65, 0.05, 0.05}, {0.12, 0.45, 0.15}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < 5; ++i) {
    vct_material& m = mats[i];
    memset(&m, 0, sizeof(m));
    for (int k = 0; k < 3; ++k) {
      double a = i >= 3 ? rng.uniform() : col[i][k];
      m.ambient[k] = (float)a;
      m.diffuse[k] = (float)a;
      m.specular[k] = i >= 3 ? 0.25f : 0.0f;
      m.transmittance[k] = 1.0f;
    }
    m.shininess = 30.0f;
    m.ior = 1.0f;
    m.dissolve = 1.0f;
    m.illum = 2;
  }
  return t == 34 ? VCT_OK : VCT_ERR_STATE;
}

}  // extern "C"
