// vct_obj.cpp - Wavefront OBJ/MTL reader behind vct_load_model.
//
// Follows what Renderer::load_model (renderer.cpp:407-603) does with the
// output of tinyobjloader (LoadObj with triangulation on, renderer.cpp:417):
//  - every shape's faces are fanned into triangles, in file order;
//  - MTL materials are appended to the material table (create_material,
//    renderer.cpp:49-81; defaults: colours 0, shininess 1, ior 1,
//    dissolve 1, illum 0);
//  - a face without material (-1) takes the file's first material, the
//    material_base_index + DEFAULT_MATERIAL_INDEX of renderer.cpp:527-537;
//  - the model dimensions use the reference's running min/max, including its
//    `max` starting at 0 and its min.z update tested against min.y
//    (renderer.cpp:457-486, 517).
// Faces without normals get the face normal (the reference leaves them 0).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/vct.h"

namespace {

std::string dir_of(const std::string& f) {
  size_t p = f.find_last_of("\\/");
  return p == std::string::npos ? std::string() : f.substr(0, p + 1);
}

// Name field of a statement: the rest of the line without the whitespace
// around it (tinyobjloader trims each line's trailing whitespace).
std::string name_of(const char* rest) {
  while (*rest == ' ' || *rest == '\t') ++rest;
  std::string n(rest);
  const size_t e = n.find_last_not_of(" \t\r\n");
  return e == std::string::npos ? std::string() : n.substr(0, e + 1);
}

void init_material(vct_material& m) {
  memset(&m, 0, sizeof(m));
  m.shininess = 1.0f;
  m.ior = 1.0f;
  m.dissolve = 1.0f;
  m.illum = 0;
}

bool parse_mtl(const std::string& path, std::vector<vct_material>& mats, std::map<std::string, int>& names) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return false;
  char line[1024];
  vct_material* cur = nullptr;
  while (fgets(line, sizeof(line), f)) {
    char* s = line;
    while (*s == ' ' || *s == '\t') ++s;
    char key[64];
    int n = 0;
    if (sscanf(s, "%63s%n", key, &n) != 1) continue;
    const char* rest = s + n;
    if (!strcmp(key, "newmtl")) {
      mats.emplace_back();
      init_material(mats.back());
      names[name_of(rest)] = (int)mats.size() - 1;
      cur = &mats.back();
      continue;
    }
    if (!cur) continue;
    float a = 0, b = 0, c = 0;
    int k = sscanf(rest, "%f %f %f", &a, &b, &c);
    auto rgb = [&](float* dst) {
      if (k >= 1) {
        dst[0] = a;
        dst[1] = k >= 3 ? b : a;
        dst[2] = k >= 3 ? c : a;
      }
    };
    if (!strcmp(key, "Ka")) rgb(cur->ambient);
    else if (!strcmp(key, "Kd")) rgb(cur->diffuse);
    else if (!strcmp(key, "Ks")) rgb(cur->specular);
    else if (!strcmp(key, "Kt") || !strcmp(key, "Tf")) rgb(cur->transmittance);
    else if (!strcmp(key, "Ke")) rgb(cur->emission);
    else if (!strcmp(key, "Ns") && k >= 1) cur->shininess = a;
    else if (!strcmp(key, "Ni") && k >= 1) cur->ior = a;
    else if (!strcmp(key, "d") && k >= 1) cur->dissolve = a;
    else if (!strcmp(key, "Tr") && k >= 1) cur->dissolve = 1.0f - a;
    else if (!strcmp(key, "illum") && k >= 1) cur->illum = (int)a;
  }
  fclose(f);
  return true;
}

struct Idx {
  int v, vn;
};

int fix_index(int i, int n) { return i > 0 ? i - 1 : (i < 0 ? n + i : -1); }

}  // namespace

// Defined in vct_runtime.hip; not part of the exported ABI.
extern "C" __attribute__((visibility("hidden"))) int vct_model_set_dims(vct_ctx ctx, uint32_t model_id,
                                                                      const float dims[3]);

// Parsed content of one OBJ file (and the MTL files it names).
struct vct_obj_data {
  std::vector<float> pos, nrm;       // 9 floats per triangle
  std::vector<int> tri_local_mat;    // index into mats, -1 = none
  std::vector<vct_material> mats;
  float dims[3] = {0, 0, 0};
};

namespace {

int parse_obj(const char* filename, vct_obj_data& o) {
  FILE* f = fopen(filename, "r");
  if (!f) return -1;
  std::vector<float> V, N;
  std::map<std::string, int> names;
  int cur_mat = -1;
  const std::string base = dir_of(filename);
  float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f}, mx[3] = {0, 0, 0};
  char line[4096];
  while (fgets(line, sizeof(line), f)) {
    char* s = line;
    while (*s == ' ' || *s == '\t') ++s;
    if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
      float a = 0, b = 0, c = 0;
      sscanf(s + 2, "%f %f %f", &a, &b, &c);
      V.push_back(a);
      V.push_back(b);
      V.push_back(c);
    } else if (s[0] == 'v' && s[1] == 'n') {
      float a = 0, b = 0, c = 0;
      sscanf(s + 3, "%f %f %f", &a, &b, &c);
      N.push_back(a);
      N.push_back(b);
      N.push_back(c);
    } else if (!strncmp(s, "mtllib", 6)) {
      parse_mtl(base + name_of(s + 6), o.mats, names);
    } else if (!strncmp(s, "usemtl", 6)) {
      auto it = names.find(name_of(s + 6));
      cur_mat = it == names.end() ? -1 : it->second;
    } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
      std::vector<Idx> face;
      char* p = s + 1;
      const int nv = (int)V.size() / 3, nn = (int)N.size() / 3;
      while (*p) {
        while (*p == ' ' || *p == '\t') ++p;
        if (!*p || *p == '\n' || *p == '\r') break;
        Idx id{-1, -1};
        id.v = fix_index((int)strtol(p, &p, 10), nv);
        if (*p == '/') {
          ++p;
          if (*p != '/') strtol(p, &p, 10);  // texcoord, unused
          if (*p == '/') {
            ++p;
            id.vn = fix_index((int)strtol(p, &p, 10), nn);
          }
        }
        while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
        if (id.v >= 0 && id.v < nv) face.push_back(id);
      }
      for (size_t k = 1; k + 1 < face.size(); ++k) {
        const Idx tri[3] = {face[0], face[k], face[k + 1]};
        float fp[3][3];
        for (int v = 0; v < 3; ++v)
          for (int c = 0; c < 3; ++c) fp[v][c] = V[tri[v].v * 3 + c];
        float e1[3], e2[3], fn[3];
        for (int c = 0; c < 3; ++c) {
          e1[c] = fp[1][c] - fp[0][c];
          e2[c] = fp[2][c] - fp[0][c];
        }
        fn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        fn[1] = e1[2] * e2[0] - e1[0] * e2[2];
        fn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        float l = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        for (int c = 0; c < 3; ++c) fn[c] = l > 0 ? fn[c] / l : 0.0f;
        for (int v = 0; v < 3; ++v) {
          for (int c = 0; c < 3; ++c) {
            o.pos.push_back(fp[v][c]);
            o.nrm.push_back(tri[v].vn >= 0 && tri[v].vn < nn ? N[tri[v].vn * 3 + c] : fn[c]);
          }
          // reference min/max bookkeeping, renderer.cpp:475-486
          if (fp[v][0] > mx[0]) mx[0] = fp[v][0];
          if (fp[v][0] < mn[0]) mn[0] = fp[v][0];
          if (fp[v][1] > mx[1]) mx[1] = fp[v][1];
          if (fp[v][1] < mn[1]) mn[1] = fp[v][1];
          if (fp[v][2] > mx[2]) mx[2] = fp[v][2];
          if (fp[v][2] < mn[1]) mn[2] = fp[v][2];
        }
        o.tri_local_mat.push_back(cur_mat);
      }
    }
  }
  fclose(f);
  for (int c = 0; c < 3; ++c) o.dims[c] = mx[c] - mn[c];
  return 0;
}

}  // namespace

extern "C" int vct_obj_open(const char* filename, vct_obj* out, uint32_t* ntri, uint32_t* nmat) {
  if (!filename || !out) return VCT_ERR_INVALID_ARG;
  *out = nullptr;
  vct_obj_data* o = new (std::nothrow) vct_obj_data;
  if (!o) return VCT_ERR_OUT_OF_MEMORY;
  if (parse_obj(filename, *o) != 0) {
    delete o;
    return VCT_ERR_NOT_FOUND;
  }
  *out = o;
  if (ntri) *ntri = (uint32_t)o->tri_local_mat.size();
  if (nmat) *nmat = (uint32_t)o->mats.size();
  return VCT_OK;
}

extern "C" int vct_obj_read(vct_obj o, float* positions, float* normals, int32_t* tri_material,
                            vct_material* materials, float dims[3]) {
  if (!o) return VCT_ERR_INVALID_ARG;
  const size_t nt = o->tri_local_mat.size();
  if (positions && nt) memcpy(positions, o->pos.data(), nt * 9 * sizeof(float));
  if (normals && nt) memcpy(normals, o->nrm.data(), nt * 9 * sizeof(float));
  if (tri_material)
    for (size_t i = 0; i < nt; ++i) tri_material[i] = o->tri_local_mat[i];
  if (materials && !o->mats.empty()) memcpy(materials, o->mats.data(), o->mats.size() * sizeof(vct_material));
  if (dims)
    for (int c = 0; c < 3; ++c) dims[c] = o->dims[c];
  return VCT_OK;
}

extern "C" void vct_obj_close(vct_obj o) { delete o; }

extern "C" int vct_load_model(vct_ctx ctx, const char* filename, uint32_t* out_id) {
  if (!out_id) return VCT_ERR_INVALID_ARG;
  *out_id = VCT_INVALID_ID;
  if (!ctx || !filename) return VCT_ERR_INVALID_ARG;
  vct_obj_data o;
  if (parse_obj(filename, o) != 0) {
    fprintf(stderr, "Error loading file %s\n", filename);  // renderer.cpp:419
    return VCT_OK;
  }
  uint32_t base_id = 0;
  if (!o.mats.empty()) {
    int rc = vct_add_materials(ctx, o.mats.data(), (uint32_t)o.mats.size(), &base_id);
    if (rc) return rc;
  }
  const uint32_t ntri = (uint32_t)o.tri_local_mat.size();
  std::vector<uint32_t> tm(ntri);
  for (uint32_t i = 0; i < ntri; ++i) {
    int lm = o.tri_local_mat[i] < 0 ? 0 : o.tri_local_mat[i];
    tm[i] = o.mats.empty() ? 0u : base_id + (uint32_t)lm;
  }
  uint32_t id = VCT_INVALID_ID;
  int rc = vct_model_create(ctx, o.pos.data(), o.nrm.data(), tm.data(), ntri, &id);
  if (rc) return rc;
  rc = vct_model_set_dims(ctx, id, o.dims);
  if (rc) return rc;
  *out_id = id;
  return VCT_OK;
}
