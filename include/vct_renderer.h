// vct_renderer.h - the reference's C++ entry points on top of the C ABI.
//
// Same names, argument meaning and call sequence as src/renderer.h:20-196 and
// src/texture_3d.h:6-10 of latencyhiding/voxel_cone_tracing, so a host
// program written against the reference switches by changing its include.
// Differences a caller sees:
//  - glm is not required: vec3/vec4/mat4 below have glm's memory layout
//    (mat4 column-major) and convert from any type with that layout through
//    the `from` helpers;
//  - load_shader is accepted and returns an id, but GLSL has no meaning here;
//  - the frame is not presented to a window: read it with image();
//  - GPU failures are reported on stderr like the reference's shader-compile
//    messages; the constructor throws when no gfx950 device is present.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "vct.h"

typedef uint32_t GLuint;
typedef float GLfloat;

namespace vct {
struct vec3 {
  float x = 0, y = 0, z = 0;
  vec3() = default;
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct vec4 {
  float x = 0, y = 0, z = 0, w = 0;
  vec4() = default;
  vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
  float& operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};
struct mat4 {
  float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // identity, column-major
  template <typename T>
  static mat4 from(const T& glm_like) {
    static_assert(sizeof(T) == 64, "needs a 4x4 float matrix");
    mat4 r;
    __builtin_memcpy(r.m, &glm_like, 64);
    return r;
  }
};
}  // namespace vct

#define INVALID_ID (~0)  // renderer.h:22
typedef size_t model_id_t;
typedef size_t shader_id_t;
typedef size_t material_id_t;

// point_light_t (renderer.h:87-92)
typedef struct {
  vct::vec3 position;
  vct::vec3 color;
  float intensity;
} point_light_t;

// material_data_t (renderer.h:94-120): packed, std140-compatible, 128 bytes
#pragma pack(push, 1)
typedef struct {
  vct::vec4 ambient;
  vct::vec4 diffuse;
  vct::vec4 specular;
  vct::vec4 transmittance;
  vct::vec3 emission;
  float shininess;
  float ior;
  float dissolve;
  int illum;
  float roughness;
  float metallic;
  float sheen;
  float clearcoat_thickness;
  float clearcoat_roughness;
  float anisotropy;
  float anisotropy_rotation;
  float pad[2];
} material_data_t;
#pragma pack(pop)

#define MAX_TEXTURE_BINDINGS 32  // renderer.h:18

// material_t (renderer.h:37-42) with the reference's members: `ubo` names the
// material (its id: there is no GL uniform buffer, the parameters live in the
// core's material table), `textures` and `num_textures` stay empty (materials
// are not texture-mapped here).  Added member: `data`, the material's
// parameters as last added or uploaded.
typedef struct {
  GLuint ubo;
  GLuint textures[MAX_TEXTURE_BINDINGS];
  size_t num_textures;
  material_data_t data;
} material_t;

struct Renderer {
  Renderer(int width, int height);  // renderer.cpp:100-147
  ~Renderer();                      // renderer.cpp:149-165

  // Resource
  model_id_t load_model(const char* filename);
  shader_id_t load_shader(const char* vertex_shader_name, const char* fragment_shader_name,
                          const char* geometry_shader_name = NULL);
  material_id_t add_material(material_data_t& material_data);
  void upload_material_data(material_data_t& material_data, material_id_t material_id);
  void set_model_material(material_id_t material_id, model_id_t model_id);

  vct::vec3 get_model_dimensions(model_id_t model);
  void set_grid_resolution(unsigned int res);
  void set_grid_size(float size);

  void set_camera_transform(vct::mat4& lookat, vct::mat4& projection);
  void set_model_transform(model_id_t model_id, vct::mat4 model_matrix);

  material_t& get_material(material_id_t material_id);

  // Render
  void set_rendering_phases(bool direct, bool diffuse, bool specular, bool shadow);
  void set_voxel_view_dir(int dir, float lod);

  void queue_model(model_id_t model_id);
  void queue_point_light(point_light_t& point_light);
  void render();

  // Additions beyond the reference's Renderer
  model_id_t add_model(const float* positions, const float* normals, const uint32_t* tri_material, uint32_t ntri);
  void set_grid_format(unsigned res, int levels, int dtype, int average);
  std::vector<float> image();  // RGBA32F, row 0 = bottom
  vct_ctx context() { return m_ctx; }

 private:
  vct_ctx m_ctx = nullptr;
  std::vector<material_t> m_materials;
  size_t m_shaders = 0;
  int m_width, m_height;
};

// Texture3D (texture_3d.h:6-10); they act on the context of the Renderer most
// recently constructed on the calling thread (as GL calls act on the current
// context).
GLuint create_tex_3d(int width, int height, int depth, int levels);
void destroy_tex_3d(GLuint tex);
void activate_tex_3d(GLuint program, GLuint tex, GLuint unit);
void clear_tex_3d(GLuint tex, GLfloat clear_color[4]);
void mip_tex_3d(GLuint tex);
