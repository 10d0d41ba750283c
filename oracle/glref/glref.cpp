// glref.cpp - runs the reference's own GLSL on the host under Mesa llvmpipe.
//
// TEST INFRASTRUCTURE ONLY (the reference oracle and the CPU baseline).
// Opens a headless OpenGL 4.5 core context through Mesa's public DRI
// software-rasteriser interface (GL/internal/dri_interface.h: dlopen of
// swrast_dri.so, DRI_SWRast.createNewScreen2, DRI_Core.createContextAttribs),
// compiles the shader files of latencyhiding/voxel_cone_tracing read from a
// directory given at run time (never copied), and replays the GL state of
// Renderer::voxelize / filter / visualize (src/renderer.cpp:283-390) with the
// textures of create_tex_3d (src/texture_3d.cpp:3-25).
//
// Deviations from the reference host code, all in state the shaders read:
//  - voxelize.frag is patched in memory, three lines: Mesa rejects a layout
//    qualifier on a function parameter (voxelize.frag:95), so the image is
//    passed by its array index instead.  The arithmetic is unchanged.
//  - glMemoryBarrier(GL_ALL_BARRIER_BITS) after the voxelize draw (the
//    reference has none before the first mip dispatch, renderer.cpp:347-350).
//  - Uniform blocks are bound at their declared bindings (camera 0,
//    material 1); the reference binds at the block index, which is the same
//    for these programs.
//  - Output goes to an RGBA32F framebuffer object instead of the window.
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#define GL_GLEXT_PROTOTYPES 0
#include <GL/glcorearb.h>
#include <GL/internal/dri_interface.h>

namespace {

std::string g_err;

#define GLF(ret, name, args) typedef ret(APIENTRY* PFN_##name) args; PFN_##name name##_ = nullptr;
GLF(void, glGenTextures, (GLsizei, GLuint*))
GLF(void, glBindTexture, (GLenum, GLuint))
GLF(void, glTexParameteri, (GLenum, GLenum, GLint))
GLF(void, glTexStorage3D, (GLenum, GLsizei, GLenum, GLsizei, GLsizei, GLsizei))
GLF(void, glTexImage3D, (GLenum, GLint, GLint, GLsizei, GLsizei, GLsizei, GLint, GLenum, GLenum, const void*))
GLF(void, glGenerateMipmap, (GLenum))
GLF(void, glDeleteTextures, (GLsizei, const GLuint*))
GLF(void, glActiveTexture, (GLenum))
GLF(void, glClearTexImage, (GLuint, GLint, GLenum, GLenum, const void*))
GLF(void, glGetTexImage, (GLenum, GLint, GLenum, GLenum, void*))
GLF(void, glTexSubImage3D, (GLenum, GLint, GLint, GLint, GLint, GLsizei, GLsizei, GLsizei, GLenum, GLenum, const void*))
GLF(GLuint, glCreateShader, (GLenum))
GLF(void, glShaderSource, (GLuint, GLsizei, const GLchar* const*, const GLint*))
GLF(void, glCompileShader, (GLuint))
GLF(void, glGetShaderiv, (GLuint, GLenum, GLint*))
GLF(void, glGetShaderInfoLog, (GLuint, GLsizei, GLsizei*, GLchar*))
GLF(GLuint, glCreateProgram, (void))
GLF(void, glAttachShader, (GLuint, GLuint))
GLF(void, glLinkProgram, (GLuint))
GLF(void, glGetProgramiv, (GLuint, GLenum, GLint*))
GLF(void, glGetProgramInfoLog, (GLuint, GLsizei, GLsizei*, GLchar*))
GLF(void, glUseProgram, (GLuint))
GLF(GLint, glGetUniformLocation, (GLuint, const GLchar*))
GLF(void, glUniform1i, (GLint, GLint))
GLF(void, glUniform1f, (GLint, GLfloat))
GLF(void, glUniform3fv, (GLint, GLsizei, const GLfloat*))
GLF(void, glUniformMatrix4fv, (GLint, GLsizei, GLboolean, const GLfloat*))
GLF(void, glGenBuffers, (GLsizei, GLuint*))
GLF(void, glBindBuffer, (GLenum, GLuint))
GLF(void, glBufferData, (GLenum, GLsizeiptr, const void*, GLenum))
GLF(void, glBindBufferBase, (GLenum, GLuint, GLuint))
GLF(void, glDeleteBuffers, (GLsizei, const GLuint*))
GLF(void, glGenVertexArrays, (GLsizei, GLuint*))
GLF(void, glBindVertexArray, (GLuint))
GLF(void, glDeleteVertexArrays, (GLsizei, const GLuint*))
GLF(void, glEnableVertexAttribArray, (GLuint))
GLF(void, glVertexAttribPointer, (GLuint, GLint, GLenum, GLboolean, GLsizei, const void*))
GLF(void, glDrawElements, (GLenum, GLsizei, GLenum, const void*))
GLF(void, glBindImageTexture, (GLuint, GLuint, GLint, GLboolean, GLint, GLenum, GLenum))
GLF(void, glDispatchCompute, (GLuint, GLuint, GLuint))
GLF(void, glMemoryBarrier, (GLbitfield))
GLF(void, glViewport, (GLint, GLint, GLsizei, GLsizei))
GLF(void, glColorMask, (GLboolean, GLboolean, GLboolean, GLboolean))
GLF(void, glDisable, (GLenum))
GLF(void, glEnable, (GLenum))
GLF(void, glBlendFunc, (GLenum, GLenum))
GLF(void, glClear, (GLbitfield))
GLF(void, glClearColor, (GLfloat, GLfloat, GLfloat, GLfloat))
GLF(void, glFinish, (void))
GLF(GLenum, glGetError, (void))
GLF(void, glGenFramebuffers, (GLsizei, GLuint*))
GLF(void, glBindFramebuffer, (GLenum, GLuint))
GLF(void, glFramebufferTexture2D, (GLenum, GLenum, GLenum, GLuint, GLint))
GLF(void, glTexImage2D, (GLenum, GLint, GLint, GLsizei, GLsizei, GLint, GLenum, GLenum, const void*))
GLF(GLenum, glCheckFramebufferStatus, (GLenum))
GLF(void, glDrawBuffers, (GLsizei, const GLenum*))
GLF(void, glReadBuffer, (GLenum))
GLF(void, glReadPixels, (GLint, GLint, GLsizei, GLsizei, GLenum, GLenum, void*))
GLF(void, glDeleteFramebuffers, (GLsizei, const GLuint*))
GLF(const GLubyte*, glGetString, (GLenum))
GLF(void, glDepthFunc, (GLenum))
GLF(void, glPixelStorei, (GLenum, GLint))

void* g_drv = nullptr;
void* g_glapi = nullptr;
const __DRIcoreExtension* g_core = nullptr;
const __DRIswrastExtension* g_swrast = nullptr;
__DRIscreen* g_screen = nullptr;
__DRIcontext* g_ctx = nullptr;
__DRIdrawable* g_draw = nullptr;
GLuint g_prog_vox = 0, g_prog_mip = 0, g_prog_vct = 0, g_prog_gbuf = 0;

void drawable_info(__DRIdrawable*, int* x, int* y, int* w, int* h, void*) {
  *x = 0;
  *y = 0;
  *w = 1;
  *h = 1;
}
void put_image(__DRIdrawable*, int, int, int, int, int, char*, void*) {}
void get_image(__DRIdrawable*, int, int, int, int, char* data, void*) {
  if (data) memset(data, 0, 4);
}

__DRIswrastLoaderExtension g_loader = {{__DRI_SWRAST_LOADER, 1}, drawable_info, put_image, get_image,
                                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
const __DRIextension* g_loader_exts[] = {&g_loader.base, nullptr};

double now_s() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

bool read_file(const std::string& path, std::string& out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[4096];
  size_t n;
  out.clear();
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  fclose(f);
  return true;
}

bool replace_once(std::string& s, const std::string& a, const std::string& b) {
  size_t p = s.find(a);
  if (p == std::string::npos) return false;
  s.replace(p, a.size(), b);
  return true;
}

GLuint compile(GLenum type, const std::string& src, const char* name) {
  GLuint sh = glCreateShader_(type);
  const char* p = src.c_str();
  glShaderSource_(sh, 1, &p, nullptr);
  glCompileShader_(sh);
  GLint ok = 0;
  glGetShaderiv_(sh, GL_COMPILE_STATUS, &ok);
  if (!ok) {
    char log[4096];
    glGetShaderInfoLog_(sh, sizeof log, nullptr, log);
    g_err = std::string("compile ") + name + ": " + log;
    return 0;
  }
  return sh;
}

GLuint link(std::vector<GLuint> shaders, const char* name) {
  GLuint pr = glCreateProgram_();
  for (GLuint s : shaders) glAttachShader_(pr, s);
  glLinkProgram_(pr);
  GLint ok = 0;
  glGetProgramiv_(pr, GL_LINK_STATUS, &ok);
  if (!ok) {
    char log[4096];
    glGetProgramInfoLog_(pr, sizeof log, nullptr, log);
    g_err = std::string("link ") + name + ": " + log;
    return 0;
  }
  return pr;
}

// G-buffer fragment stage (project code) behind the reference vertex shader.
const char* kGbufFrag = R"(#version 450 core
in VS_OUT { vec4 world_position; vec3 normal; vec2 uv; } vs_out;
uniform uint material_id;
layout (location = 0) out vec4 out_pos;
layout (location = 1) out vec4 out_nrm;
void main() {
  out_pos = vec4(vs_out.world_position.xyz, 1.0 / gl_FragCoord.w);
  out_nrm = vec4(vs_out.normal, uintBitsToFloat(material_id));
}
)";

struct Scene {
  GLuint vao = 0, vbo = 0, ebo = 0;
  std::vector<GLuint> ubos;                         // one per material
  std::vector<std::pair<uint32_t, uint32_t>> runs;  // (first triangle, count) per draw object
  std::vector<uint32_t> run_mat;
};

// Geometry of one model: vert_data_t {pos3, norm3, tex2} (renderer.cpp:23-35),
// one index per vertex, one draw object per run of equal material.
void build_scene(Scene& sc, int ntri, const float* pos, const float* nrm, const uint32_t* tri_mat,
                 const void* mats, int nmat) {
  std::vector<float> v((size_t)ntri * 3 * 8, 0.0f);
  std::vector<uint32_t> idx((size_t)ntri * 3);
  for (size_t i = 0; i < (size_t)ntri * 3; ++i) {
    for (int k = 0; k < 3; ++k) {
      v[i * 8 + k] = pos[i * 3 + k];
      v[i * 8 + 3 + k] = nrm[i * 3 + k];
    }
    idx[i] = (uint32_t)i;
  }
  glGenVertexArrays_(1, &sc.vao);
  glGenBuffers_(1, &sc.vbo);
  glGenBuffers_(1, &sc.ebo);
  glBindVertexArray_(sc.vao);
  glBindBuffer_(GL_ARRAY_BUFFER, sc.vbo);
  glBufferData_(GL_ARRAY_BUFFER, v.size() * 4, v.data(), GL_STATIC_DRAW);
  glBindBuffer_(GL_ELEMENT_ARRAY_BUFFER, sc.ebo);
  glBufferData_(GL_ELEMENT_ARRAY_BUFFER, idx.size() * 4, idx.data(), GL_STATIC_DRAW);
  glEnableVertexAttribArray_(0);
  glVertexAttribPointer_(0, 3, GL_FLOAT, GL_FALSE, 32, (const void*)0);
  glEnableVertexAttribArray_(1);
  glVertexAttribPointer_(1, 3, GL_FLOAT, GL_FALSE, 32, (const void*)12);
  glEnableVertexAttribArray_(2);
  glVertexAttribPointer_(2, 2, GL_FLOAT, GL_FALSE, 32, (const void*)24);
  glBindVertexArray_(0);
  sc.ubos.resize(nmat);
  glGenBuffers_(nmat, sc.ubos.data());
  for (int m = 0; m < nmat; ++m) {
    glBindBuffer_(GL_UNIFORM_BUFFER, sc.ubos[m]);
    glBufferData_(GL_UNIFORM_BUFFER, 128, (const char*)mats + 128 * m, GL_STATIC_DRAW);
  }
  glBindBuffer_(GL_UNIFORM_BUFFER, 0);
  for (int t = 0; t < ntri;) {
    int e = t;
    const uint32_t m = tri_mat ? tri_mat[t] : 0;
    while (e < ntri && (tri_mat ? tri_mat[e] : 0) == m) ++e;
    sc.runs.push_back({(uint32_t)t, (uint32_t)(e - t)});
    sc.run_mat.push_back(m);
    t = e;
  }
}

void free_scene(Scene& sc) {
  glDeleteVertexArrays_(1, &sc.vao);
  glDeleteBuffers_(1, &sc.vbo);
  glDeleteBuffers_(1, &sc.ebo);
  if (!sc.ubos.empty()) glDeleteBuffers_((GLsizei)sc.ubos.size(), sc.ubos.data());
}

// draw_models (renderer.cpp:240-257) for our single model
void draw(GLuint prog, const Scene& sc, const float* model, bool set_mat_uniform) {
  glUniformMatrix4fv_(glGetUniformLocation_(prog, "model"), 1, GL_FALSE, model);
  glBindVertexArray_(sc.vao);
  const GLint mloc = set_mat_uniform ? glGetUniformLocation_(prog, "material_id") : -1;
  for (size_t r = 0; r < sc.runs.size(); ++r) {
    glBindBufferBase_(GL_UNIFORM_BUFFER, 1, sc.ubos[sc.run_mat[r]]);
    if (mloc >= 0) {
      typedef void(APIENTRY * PFN_u)(GLint, GLuint);
      static PFN_u u1ui = nullptr;
      if (!u1ui) u1ui = (PFN_u)((void* (*)(const char*))dlsym(g_glapi, "_glapi_get_proc_address"))("glUniform1ui");
      u1ui(mloc, sc.run_mat[r]);
    }
    glDrawElements_(GL_TRIANGLES, (GLsizei)(sc.runs[r].second * 3), GL_UNSIGNED_INT,
                    (const void*)(uintptr_t)(sc.runs[r].first * 3 * 4));
  }
  glBindVertexArray_(0);
}

void upload_lights(GLuint prog, const float* lights, int nlights) {
  for (int i = 0; i < nlights; ++i) {
    char n[64];
    snprintf(n, sizeof n, "point_lights[%d].position", i);
    glUniform3fv_(glGetUniformLocation_(prog, n), 1, lights + i * 8);
    snprintf(n, sizeof n, "point_lights[%d].color", i);
    glUniform3fv_(glGetUniformLocation_(prog, n), 1, lights + i * 8 + 3);
    snprintf(n, sizeof n, "point_lights[%d].intensity", i);
    glUniform1f_(glGetUniformLocation_(prog, n), lights[i * 8 + 6]);
  }
  glUniform1i_(glGetUniformLocation_(prog, "point_light_count"), nlights);
}

GLuint camera_ubo(const float* view, const float* proj) {
  float cam[32];
  memcpy(cam, proj, 64);
  memcpy(cam + 16, view, 64);
  GLuint u = 0;
  glGenBuffers_(1, &u);
  glBindBuffer_(GL_UNIFORM_BUFFER, u);
  glBufferData_(GL_UNIFORM_BUFFER, 128, cam, GL_DYNAMIC_DRAW);
  glBindBuffer_(GL_UNIFORM_BUFFER, 0);
  return u;
}

// create_tex_3d (texture_3d.cpp:3-25), including its ignored calls
GLuint create_tex_3d(int w, int h, int d, int levels) {
  GLuint tex;
  glGenTextures_(1, &tex);
  glBindTexture_(GL_TEXTURE_3D, tex);
  glTexParameteri_(GL_TEXTURE_3D, GL_TEXTURE_WRAP_S, GL_CLAMP_TO_BORDER);
  glTexParameteri_(GL_TEXTURE_3D, GL_TEXTURE_WRAP_T, GL_CLAMP_TO_BORDER);
  glTexParameteri_(GL_TEXTURE_3D, GL_TEXTURE_WRAP_R, GL_CLAMP_TO_BORDER);
  glTexParameteri_(GL_TEXTURE_3D, GL_TEXTURE_MIN_FILTER, GL_LINEAR_MIPMAP_LINEAR);
  glTexParameteri_(GL_TEXTURE_3D, GL_TEXTURE_MAG_FILTER, GL_LINEAR_MIPMAP_LINEAR);  // invalid: ignored
  glTexStorage3D_(GL_TEXTURE_3D, levels, GL_RGBA8, w, h, d);
  glTexImage3D_(GL_TEXTURE_3D, 0, GL_R32UI, w, h, d, 0, GL_RGBA, GL_FLOAT, nullptr);  // immutable: error
  glGenerateMipmap_(GL_TEXTURE_3D);
  glBindTexture_(GL_TEXTURE_3D, 0);
  while (glGetError_() != GL_NO_ERROR) {
  }
  return tex;
}

}  // namespace

extern "C" {

const char* glref_error(void) { return g_err.c_str(); }

const char* glref_renderer(void) {
  if (!glGetString_) return "";
  return (const char*)glGetString_(GL_RENDERER);
}

// Open the context and build the three reference programs from shader_dir.
int glref_init(const char* shader_dir) {
  if (g_ctx) return 0;
  const char* drv_paths[] = {"/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so", "swrast_dri.so"};
  for (const char* p : drv_paths)
    if (!g_drv) g_drv = dlopen(p, RTLD_NOW | RTLD_GLOBAL);
  g_glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
  if (!g_drv || !g_glapi) {
    g_err = "Mesa swrast_dri.so / libglapi.so.0 not found";
    return -1;
  }
  typedef const __DRIextension** (*GetExt)(void);
  GetExt get = (GetExt)dlsym(g_drv, "__driDriverGetExtensions_swrast");
  if (!get) {
    g_err = "__driDriverGetExtensions_swrast missing";
    return -1;
  }
  const __DRIextension** exts = get();
  for (int i = 0; exts[i]; ++i) {
    if (!strcmp(exts[i]->name, __DRI_CORE)) g_core = (const __DRIcoreExtension*)exts[i];
    if (!strcmp(exts[i]->name, __DRI_SWRAST)) g_swrast = (const __DRIswrastExtension*)exts[i];
  }
  if (!g_core || !g_swrast || g_swrast->base.version < 4) {
    g_err = "DRI core/swrast extensions missing";
    return -1;
  }
  const __DRIconfig** configs = nullptr;
  g_screen = g_swrast->createNewScreen2(0, g_loader_exts, exts, &configs, nullptr);
  if (!g_screen || !configs || !configs[0]) {
    g_err = "createNewScreen2 failed";
    return -1;
  }
  const uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 5,
                              __DRI_CTX_ATTRIB_FLAGS, __DRI_CTX_FLAG_FORWARD_COMPATIBLE};
  unsigned err = 0;
  g_ctx = g_swrast->createContextAttribs(g_screen, __DRI_API_OPENGL_CORE, configs[0], nullptr, 3, attribs, &err,
                                         nullptr);
  if (!g_ctx) {
    g_err = "createContextAttribs(4.5 core) failed";
    return -1;
  }
  g_draw = g_swrast->createNewDrawable(g_screen, configs[0], nullptr);
  if (!g_core->bindContext(g_ctx, g_draw, g_draw)) {
    g_err = "bindContext failed";
    return -1;
  }
  typedef void* (*GPA)(const char*);
  GPA gpa = (GPA)dlsym(g_glapi, "_glapi_get_proc_address");
#define LOAD(name)                                  \
  name##_ = (PFN_##name)gpa(#name);                 \
  if (!name##_) {                                   \
    g_err = "missing GL entry point " #name;        \
    return -1;                                      \
  }
  LOAD(glGenTextures) LOAD(glBindTexture) LOAD(glTexParameteri) LOAD(glTexStorage3D) LOAD(glTexImage3D)
  LOAD(glGenerateMipmap) LOAD(glDeleteTextures) LOAD(glActiveTexture) LOAD(glClearTexImage) LOAD(glGetTexImage)
  LOAD(glTexSubImage3D) LOAD(glCreateShader) LOAD(glShaderSource) LOAD(glCompileShader) LOAD(glGetShaderiv)
  LOAD(glGetShaderInfoLog) LOAD(glCreateProgram) LOAD(glAttachShader) LOAD(glLinkProgram) LOAD(glGetProgramiv)
  LOAD(glGetProgramInfoLog) LOAD(glUseProgram) LOAD(glGetUniformLocation) LOAD(glUniform1i) LOAD(glUniform1f)
  LOAD(glUniform3fv) LOAD(glUniformMatrix4fv) LOAD(glGenBuffers) LOAD(glBindBuffer) LOAD(glBufferData)
  LOAD(glBindBufferBase) LOAD(glDeleteBuffers) LOAD(glGenVertexArrays) LOAD(glBindVertexArray)
  LOAD(glDeleteVertexArrays) LOAD(glEnableVertexAttribArray) LOAD(glVertexAttribPointer) LOAD(glDrawElements)
  LOAD(glBindImageTexture) LOAD(glDispatchCompute) LOAD(glMemoryBarrier) LOAD(glViewport) LOAD(glColorMask)
  LOAD(glDisable) LOAD(glEnable) LOAD(glBlendFunc) LOAD(glClear) LOAD(glClearColor) LOAD(glFinish)
  LOAD(glGetError) LOAD(glGenFramebuffers) LOAD(glBindFramebuffer) LOAD(glFramebufferTexture2D)
  LOAD(glTexImage2D) LOAD(glCheckFramebufferStatus) LOAD(glDrawBuffers) LOAD(glReadBuffer) LOAD(glReadPixels)
  LOAD(glDeleteFramebuffers) LOAD(glGetString) LOAD(glDepthFunc) LOAD(glPixelStorei)
#undef LOAD
  const std::string dir = shader_dir;
  std::string vv, vg, vf, mc, tv, tf;
  if (!read_file(dir + "/voxelize.vert", vv) || !read_file(dir + "/voxelize.geom", vg) ||
      !read_file(dir + "/voxelize.frag", vf) || !read_file(dir + "/mipmap.comp", mc) ||
      !read_file(dir + "/voxel_cone_tracing.vert", tv) || !read_file(dir + "/voxel_cone_tracing.frag", tf)) {
    g_err = "cannot read the reference shaders in " + dir;
    return -1;
  }
  // voxelize.frag:95,106,160 - pass the image index instead of the image
  bool ok = replace_once(vf, "void imageAtomicRGBA8Avg(layout (r32ui) uimage3D img, ivec3 coords, vec4 val)",
                         "void imageAtomicRGBA8Avg(int img, ivec3 coords, vec4 val)") &&
            replace_once(vf, "imageAtomicCompSwap(img, coords, prevStoredVal, newVal)",
                         "imageAtomicCompSwap(tex3D[img], coords, prevStoredVal, newVal)") &&
            replace_once(vf, "imageAtomicRGBA8Avg(tex3D[i], voxel_pos, final_color);",
                         "imageAtomicRGBA8Avg(i, voxel_pos, final_color);");
  if (!ok) {
    g_err = "voxelize.frag does not have the expected lines";
    return -1;
  }
  GLuint s1 = compile(GL_VERTEX_SHADER, vv, "voxelize.vert"), s2 = compile(GL_GEOMETRY_SHADER, vg, "voxelize.geom"),
         s3 = compile(GL_FRAGMENT_SHADER, vf, "voxelize.frag");
  if (!s1 || !s2 || !s3 || !(g_prog_vox = link({s1, s2, s3}, "voxelize"))) return -1;
  GLuint c1 = compile(GL_COMPUTE_SHADER, mc, "mipmap.comp");
  if (!c1 || !(g_prog_mip = link({c1}, "mipmap"))) return -1;
  GLuint t1 = compile(GL_VERTEX_SHADER, tv, "voxel_cone_tracing.vert"),
         t2 = compile(GL_FRAGMENT_SHADER, tf, "voxel_cone_tracing.frag");
  if (!t1 || !t2 || !(g_prog_vct = link({t1, t2}, "voxel_cone_tracing"))) return -1;
  GLuint t3 = compile(GL_VERTEX_SHADER, tv, "voxel_cone_tracing.vert"),
         g1 = compile(GL_FRAGMENT_SHADER, kGbufFrag, "gbuffer.frag");
  if (!t3 || !g1 || !(g_prog_gbuf = link({t3, g1}, "gbuffer"))) return -1;
  return 0;
}

// One reference frame.  Inputs: res (7 levels, as set_grid_resolution), cube
// size, model matrix, camera (column-major view, projection), triangles
// (9 floats per triangle for positions and normals), material per triangle,
// 128-byte material records, point lights (8 floats: position, colour,
// intensity, unused), render flags.  Outputs (each may be NULL):
//   level0  res^3 RGBA8 texels of map 0 (x fastest)
//   mips    levels 1..6 of the 6 maps, (res>>l)^3 RGBA8 each, level-major
//   image   W x H RGBA32F of voxel_cone_tracing.frag (row 0 = bottom)
//   gpos/gnrm W x H RGBA32F G-buffer (world xyz, 1/gl_FragCoord.w; normal, material)
//   times   [voxelize, mip, trace] seconds (glFinish-fenced)
int glref_frame(int res, float cube_size, const float* model, const float* view, const float* proj, int W, int H,
                int ntri, const float* pos, const float* nrm, const uint32_t* tri_mat, const void* mats, int nmat,
                const float* lights, int nlights, int direct, int diffuse, int specular, int shadow,
                int view_voxel_dir, float view_voxel_lod, uint32_t* level0, uint32_t* mips, float* image,
                float* gpos, float* gnrm, double* times) {
  if (!g_ctx) {
    g_err = "glref_init not called";
    return -1;
  }
  while (glGetError_() != GL_NO_ERROR) {
  }
  Scene sc;
  build_scene(sc, ntri, pos, nrm, tri_mat, mats, nmat);
  GLuint maps[6];
  for (int i = 0; i < 6; ++i) maps[i] = create_tex_3d(res, res, res, 7);
  GLuint cam = camera_ubo(view, proj);
  const float cam_pos[3] = {view[12], view[13], view[14]};  // renderer.cpp:279
  // ---- voxelize (renderer.cpp:316-348) ----
  // The reference draws into its 800x600 window; here an off-screen target of
  // the voxelize viewport's size keeps every fragment inside the framebuffer.
  GLuint vfbo = 0, vcol = 0;
  glGenFramebuffers_(1, &vfbo);
  glBindFramebuffer_(GL_FRAMEBUFFER, vfbo);
  glGenTextures_(1, &vcol);
  glBindTexture_(GL_TEXTURE_2D, vcol);
  glTexImage2D_(GL_TEXTURE_2D, 0, GL_R8, 2 * res, 2 * res, 0, GL_RED, GL_UNSIGNED_BYTE, nullptr);
  glFramebufferTexture2D_(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT0, GL_TEXTURE_2D, vcol, 0);
  if (glCheckFramebufferStatus_(GL_FRAMEBUFFER) != GL_FRAMEBUFFER_COMPLETE) {
    g_err = "voxelize framebuffer incomplete";
    return -1;
  }
  glFinish_();
  double t0 = now_s();
  const float black[4] = {0, 0, 0, 0};
  for (int i = 0; i < 6; ++i) glClearTexImage_(maps[i], 0, GL_RGBA, GL_FLOAT, black);
  glUseProgram_(g_prog_vox);
  glUniform1f_(glGetUniformLocation_(g_prog_vox, "cube_size"), cube_size);
  for (int i = 0; i < 6; ++i) glBindImageTexture_(i + 2, maps[i], 0, GL_TRUE, 0, GL_WRITE_ONLY, GL_R32UI);
  glBindBufferBase_(GL_UNIFORM_BUFFER, 0, cam);
  glUniform3fv_(glGetUniformLocation_(g_prog_vox, "camera_position"), 1, cam_pos);
  upload_lights(g_prog_vox, lights, nlights);
  glViewport_(0, 0, 2 * res, 2 * res);
  glColorMask_(GL_FALSE, GL_FALSE, GL_FALSE, GL_FALSE);
  glDisable_(GL_CULL_FACE);
  glDisable_(GL_DEPTH_TEST);
  draw(g_prog_vox, sc, model, false);
  glMemoryBarrier_(GL_ALL_BARRIER_BITS);
  glFinish_();
  double t1 = now_s();
  // ---- filter (renderer.cpp:283-314) ----
  glUseProgram_(g_prog_mip);
  for (int i = 0; i < 6; ++i) {
    glActiveTexture_(GL_TEXTURE0 + i);
    glBindTexture_(GL_TEXTURE_3D, maps[i]);
  }
  glActiveTexture_(GL_TEXTURE0);
  const GLint rl = glGetUniformLocation_(g_prog_mip, "resolution"), ml = glGetUniformLocation_(g_prog_mip, "mip");
  int mip = 0;
  for (size_t d = (size_t)res; d >= 1; d /= 2) {
    glUniform1i_(rl, (GLint)d);
    glUniform1i_(ml, mip);
    for (int i = 0; i < 6; ++i) glBindImageTexture_(i, maps[i], mip + 1, GL_TRUE, 0, GL_WRITE_ONLY, GL_RGBA8);
    const unsigned wg = (unsigned)((d + 7) / 8);
    glDispatchCompute_(wg, wg, wg);
    glMemoryBarrier_(GL_TEXTURE_FETCH_BARRIER_BIT | GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
    ++mip;
  }
  glFinish_();
  double t2 = now_s();
  while (glGetError_() != GL_NO_ERROR) {  // binding non-existent levels >= 7 is an error, as in the reference
  }
  glColorMask_(GL_TRUE, GL_TRUE, GL_TRUE, GL_TRUE);
  // ---- visualize (renderer.cpp:355-390) into an RGBA32F target ----
  GLuint fbo = 0, col = 0, dep = 0, gp = 0, gn = 0;
  glGenFramebuffers_(1, &fbo);
  glBindFramebuffer_(GL_FRAMEBUFFER, fbo);
  glGenTextures_(1, &col);
  glBindTexture_(GL_TEXTURE_2D, col);
  glTexImage2D_(GL_TEXTURE_2D, 0, GL_RGBA32F, W, H, 0, GL_RGBA, GL_FLOAT, nullptr);
  glFramebufferTexture2D_(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT0, GL_TEXTURE_2D, col, 0);
  glGenTextures_(1, &dep);
  glBindTexture_(GL_TEXTURE_2D, dep);
  glTexImage2D_(GL_TEXTURE_2D, 0, GL_DEPTH_COMPONENT32F, W, H, 0, GL_DEPTH_COMPONENT, GL_FLOAT, nullptr);
  glFramebufferTexture2D_(GL_FRAMEBUFFER, GL_DEPTH_ATTACHMENT, GL_TEXTURE_2D, dep, 0);
  if (glCheckFramebufferStatus_(GL_FRAMEBUFFER) != GL_FRAMEBUFFER_COMPLETE) {
    g_err = "framebuffer incomplete";
    return -1;
  }
  glViewport_(0, 0, W, H);
  glEnable_(GL_DEPTH_TEST);
  glClearColor_(0.15f, 0.25f, 0.25f, 1.0f);
  glClear_(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
  glUseProgram_(g_prog_vct);
  glUniform1f_(glGetUniformLocation_(g_prog_vct, "cube_size"), cube_size);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "cube_res"), res);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "enable_diffuse"), diffuse);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "enable_specular"), specular);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "enable_shadow"), shadow);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "enable_direct"), direct);
  glUniform1i_(glGetUniformLocation_(g_prog_vct, "view_voxel_dir"), view_voxel_dir);
  glUniform1f_(glGetUniformLocation_(g_prog_vct, "view_voxel_lod"), view_voxel_lod);
  for (int i = 0; i < 6; ++i) {
    glActiveTexture_(GL_TEXTURE0 + 2 + i);
    glBindTexture_(GL_TEXTURE_3D, maps[i]);
  }
  glBindBufferBase_(GL_UNIFORM_BUFFER, 0, cam);
  glUniform3fv_(glGetUniformLocation_(g_prog_vct, "camera_position"), 1, cam_pos);
  upload_lights(g_prog_vct, lights, nlights);
  glEnable_(GL_BLEND);
  glBlendFunc_(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA);
  glFinish_();
  double t3 = now_s();
  draw(g_prog_vct, sc, model, false);
  glFinish_();
  double t4 = now_s();
  glDisable_(GL_BLEND);
  glPixelStorei_(GL_PACK_ALIGNMENT, 1);
  if (image) {
    glReadBuffer_(GL_COLOR_ATTACHMENT0);
    glReadPixels_(0, 0, W, H, GL_RGBA, GL_FLOAT, image);
  }
  if (gpos || gnrm) {
    glGenTextures_(1, &gp);
    glBindTexture_(GL_TEXTURE_2D, gp);
    glTexImage2D_(GL_TEXTURE_2D, 0, GL_RGBA32F, W, H, 0, GL_RGBA, GL_FLOAT, nullptr);
    glGenTextures_(1, &gn);
    glBindTexture_(GL_TEXTURE_2D, gn);
    glTexImage2D_(GL_TEXTURE_2D, 0, GL_RGBA32F, W, H, 0, GL_RGBA, GL_FLOAT, nullptr);
    glFramebufferTexture2D_(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT0, GL_TEXTURE_2D, gp, 0);
    glFramebufferTexture2D_(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT1, GL_TEXTURE_2D, gn, 0);
    const GLenum bufs[2] = {GL_COLOR_ATTACHMENT0, GL_COLOR_ATTACHMENT1};
    glDrawBuffers_(2, bufs);
    glClearColor_(0, 0, 0, 0);
    glClear_(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
    glUseProgram_(g_prog_gbuf);
    glBindBufferBase_(GL_UNIFORM_BUFFER, 0, cam);
    draw(g_prog_gbuf, sc, model, true);
    glFinish_();
    if (gpos) {
      glReadBuffer_(GL_COLOR_ATTACHMENT0);
      glReadPixels_(0, 0, W, H, GL_RGBA, GL_FLOAT, gpos);
    }
    if (gnrm) {
      glReadBuffer_(GL_COLOR_ATTACHMENT1);
      glReadPixels_(0, 0, W, H, GL_RGBA, GL_FLOAT, gnrm);
    }
  }
  // ---- read back the maps ----
  glPixelStorei_(GL_PACK_ALIGNMENT, 1);
  if (level0) {
    glBindTexture_(GL_TEXTURE_3D, maps[0]);
    glGetTexImage_(GL_TEXTURE_3D, 0, GL_RGBA, GL_UNSIGNED_BYTE, level0);
  }
  if (mips) {
    size_t off = 0;
    for (int l = 1; l < 7; ++l) {
      const size_t s = (size_t)(res >> l);
      if (s == 0) break;
      for (int d = 0; d < 6; ++d) {
        glBindTexture_(GL_TEXTURE_3D, maps[d]);
        glGetTexImage_(GL_TEXTURE_3D, l, GL_RGBA, GL_UNSIGNED_BYTE, mips + off);
        off += s * s * s;
      }
    }
  }
  glBindTexture_(GL_TEXTURE_3D, 0);
  if (times) {
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t4 - t3;
  }
  glBindFramebuffer_(GL_FRAMEBUFFER, 0);
  glDeleteFramebuffers_(1, &fbo);
  glDeleteFramebuffers_(1, &vfbo);
  GLuint texs[5] = {col, dep, gp, gn, vcol};
  glDeleteTextures_(5, texs);
  glDeleteTextures_(6, maps);
  glDeleteBuffers_(1, &cam);
  free_scene(sc);
  const GLenum e = glGetError_();
  if (e != GL_NO_ERROR) {
    char b[64];
    snprintf(b, sizeof b, "GL error 0x%x", e);
    g_err = b;
    return -2;
  }
  return 0;
}

}  // extern "C"
