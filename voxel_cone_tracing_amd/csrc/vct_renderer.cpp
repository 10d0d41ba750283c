// vct_renderer.cpp - Renderer / Texture3D (include/vct_renderer.h) over the C ABI.
#include "../../include/vct_renderer.h"

#include <stdio.h>
#include <string.h>

#include <stdexcept>
#include <string>

namespace {

thread_local vct_ctx t_current = nullptr;

void report(const char* what, int rc) {
  if (rc != VCT_OK) fprintf(stderr, "%s failed (%d): %s\n", what, rc, vct_last_error());
}

vct_material to_c(const material_data_t& d) {
  static_assert(sizeof(material_data_t) == sizeof(vct_material), "material_data_t layout");
  vct_material m;
  memcpy(&m, &d, sizeof m);
  return m;
}

// fill_default_mat_data (renderer.cpp:85-98)
material_data_t default_material() {
  material_data_t m;
  memset(&m, 0, sizeof m);
  m.ambient = vct::vec4(1, 0, 1, 0);
  m.diffuse = vct::vec4(1, 0, 1, 0);
  m.shininess = 1;
  m.ior = 1;
  m.dissolve = 1;
  m.illum = 0;
  return m;
}

}  // namespace

// A table entry of material `id` with parameters d.
static material_t material_entry(size_t id, const material_data_t& d) {
  material_t mt;
  memset(&mt, 0, sizeof mt);
  mt.ubo = (GLuint)id;
  mt.num_textures = 0;
  mt.data = d;
  return mt;
}

Renderer::Renderer(int width, int height) : m_width(width), m_height(height) {
  int rc = vct_create(width, height, &m_ctx);
  if (rc != VCT_OK) throw std::runtime_error(std::string("Renderer: ") + vct_last_error());
  m_materials.push_back(material_entry(0, default_material()));  // material 0, as in renderer.cpp:115-128
  t_current = m_ctx;
  set_grid_resolution(128);  // renderer.cpp:146
}

Renderer::~Renderer() {
  if (t_current == m_ctx) t_current = nullptr;
  vct_destroy(m_ctx);
}

model_id_t Renderer::load_model(const char* filename) {
  uint32_t id = VCT_INVALID_ID;
  int rc = vct_load_model(m_ctx, filename, &id);
  report("load_model", rc);
  if (rc != VCT_OK || id == VCT_INVALID_ID) return (model_id_t)INVALID_ID;
  // mirror the materials the loader appended
  vct_material m;
  for (size_t i = m_materials.size(); vct_get_material(m_ctx, (uint32_t)i, &m) == VCT_OK; ++i) {
    material_data_t d;
    memcpy(&d, &m, sizeof m);
    m_materials.push_back(material_entry(i, d));
  }
  return id;
}

shader_id_t Renderer::load_shader(const char*, const char*, const char*) { return m_shaders++; }

material_id_t Renderer::add_material(material_data_t& material_data) {
  vct_material m = to_c(material_data);
  uint32_t id = 0;
  report("add_material", vct_add_material(m_ctx, &m, &id));
  m_materials.push_back(material_entry(id, material_data));
  return id;
}

void Renderer::upload_material_data(material_data_t& material_data, material_id_t material_id) {
  vct_material m = to_c(material_data);
  report("upload_material_data", vct_upload_material_data(m_ctx, &m, (uint32_t)material_id));
  if (material_id < m_materials.size()) m_materials[material_id].data = material_data;
}

void Renderer::set_model_material(material_id_t material_id, model_id_t model_id) {
  report("set_model_material", vct_set_model_material(m_ctx, (uint32_t)material_id, (uint32_t)model_id));
}

vct::vec3 Renderer::get_model_dimensions(model_id_t model) {
  float d[3] = {0, 0, 0};
  report("get_model_dimensions", vct_get_model_dimensions(m_ctx, (uint32_t)model, d));
  return vct::vec3(d[0], d[1], d[2]);
}

void Renderer::set_grid_resolution(unsigned int res) {
  report("set_grid_resolution", vct_set_grid_resolution(m_ctx, res));
}

void Renderer::set_grid_format(unsigned res, int levels, int dtype, int average) {
  report("set_grid_format", vct_set_grid_format(m_ctx, res, levels, dtype, average));
}

void Renderer::set_grid_size(float size) { report("set_grid_size", vct_set_grid_size(m_ctx, size)); }

void Renderer::set_camera_transform(vct::mat4& lookat, vct::mat4& projection) {
  report("set_camera_transform", vct_set_camera_transform(m_ctx, lookat.m, projection.m));
}

void Renderer::set_model_transform(model_id_t model_id, vct::mat4 model_matrix) {
  report("set_model_transform", vct_set_model_transform(m_ctx, (uint32_t)model_id, model_matrix.m));
}

material_t& Renderer::get_material(material_id_t material_id) { return m_materials.at(material_id); }

void Renderer::set_rendering_phases(bool direct, bool diffuse, bool specular, bool shadow) {
  report("set_rendering_phases", vct_set_rendering_phases(m_ctx, direct, diffuse, specular, shadow));
}

void Renderer::set_voxel_view_dir(int dir, float lod) {
  report("set_voxel_view_dir", vct_set_voxel_view_dir(m_ctx, dir, lod));
}

void Renderer::queue_model(model_id_t model_id) {
  if (model_id > 0xFFFFFFFFull) return;  // ignored like renderer.cpp:211-212
  report("queue_model", vct_queue_model(m_ctx, (uint32_t)model_id));
}

void Renderer::queue_point_light(point_light_t& point_light) {
  vct_light l;
  l.position[0] = point_light.position.x;
  l.position[1] = point_light.position.y;
  l.position[2] = point_light.position.z;
  l.color[0] = point_light.color.x;
  l.color[1] = point_light.color.y;
  l.color[2] = point_light.color.z;
  l.intensity = point_light.intensity;
  l.kind = VCT_LIGHT_POINT;
  report("queue_point_light", vct_queue_point_light(m_ctx, &l));
}

void Renderer::render() { report("render", vct_render(m_ctx)); }

model_id_t Renderer::add_model(const float* positions, const float* normals, const uint32_t* tri_material,
                               uint32_t ntri) {
  uint32_t id = VCT_INVALID_ID;
  report("add_model", vct_model_create(m_ctx, positions, normals, tri_material, ntri, &id));
  return id == VCT_INVALID_ID ? (model_id_t)INVALID_ID : id;
}

std::vector<float> Renderer::image() {
  std::vector<float> img((size_t)m_width * m_height * 4, 0.0f);
  report("image", vct_image_download(m_ctx, img.data(), (size_t)m_width * m_height));
  return img;
}

// ---- Texture3D ------------------------------------------------------------------
GLuint create_tex_3d(int width, int height, int depth, int levels) {
  uint32_t t = 0;
  report("create_tex_3d", vct_tex3d_create(t_current, width, height, depth, levels, VCT_RGBA8, &t));
  return t;
}

void destroy_tex_3d(GLuint tex) { report("destroy_tex_3d", vct_tex3d_destroy(t_current, tex)); }

void activate_tex_3d(GLuint, GLuint tex, GLuint unit) {
  report("activate_tex_3d", vct_tex3d_activate(t_current, tex, unit));
}

void clear_tex_3d(GLuint tex, GLfloat clear_color[4]) {
  report("clear_tex_3d", vct_tex3d_clear(t_current, tex, clear_color));
}

void mip_tex_3d(GLuint tex) { report("mip_tex_3d", vct_tex3d_mip(t_current, tex)); }
