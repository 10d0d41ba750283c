// vct_reference_app.cpp - the reference application's scene and frame loop
// (src/main.cpp:85-121 and 141-388) on the Renderer drop-in, without window,
// UI or input: the Cornell box and suzanne loaded with load_model, suzanne's
// material (main.cpp:89-100) set with add_material / set_model_material and
// re-uploaded every frame, grid size 3 at the Renderer's default 128^3, the
// camera at (0, 0.9, 3) with pitch 0 and yaw -90 (camera.h:26-58) and
// glm::perspective(45 [radians], 4/3, 0.1, 100), the ceiling light, every
// rendering phase on, suzanne at translate(0, 1.1, -0.5) * rotate(0) *
// scale(0.3).  Writes the last frame's image (RGBA32F, row 0 at the bottom).
//
//   vct_reference_app BOX.obj SUZANNE.obj WIDTH HEIGHT OUT.f32 [FRAMES]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/vct_renderer.h"

namespace {

struct V3 {
  float x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalize(V3 a) {
  const float l = sqrtf(dot(a, a));
  return {a.x / l, a.y / l, a.z / l};
}

// glm::lookAt (right-handed), column-major
vct::mat4 look_at(V3 eye, V3 center, V3 up) {
  const V3 f = normalize(sub(center, eye));
  const V3 s = normalize(cross(f, up));
  const V3 u = cross(s, f);
  vct::mat4 m;
  float* M = m.m;
  M[0] = s.x, M[4] = s.y, M[8] = s.z;
  M[1] = u.x, M[5] = u.y, M[9] = u.z;
  M[2] = -f.x, M[6] = -f.y, M[10] = -f.z;
  M[3] = M[7] = M[11] = 0.0f;
  M[12] = -dot(s, eye);
  M[13] = -dot(u, eye);
  M[14] = dot(f, eye);
  M[15] = 1.0f;
  return m;
}

// glm::perspective (right-handed, clip z in [-w, w]); fovy in radians
vct::mat4 perspective(float fovy, float aspect, float n, float f) {
  vct::mat4 m;
  float* M = m.m;
  for (int i = 0; i < 16; ++i) M[i] = 0.0f;
  const float t = tanf(fovy / 2.0f);
  M[0] = 1.0f / (aspect * t);
  M[5] = 1.0f / t;
  M[10] = -(f + n) / (f - n);
  M[11] = -1.0f;
  M[14] = -(2.0f * f * n) / (f - n);
  return m;
}

float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s BOX.obj SUZANNE.obj WIDTH HEIGHT OUT.f32 [FRAMES]\n", argv[0]);
    return 2;
  }
  const int W = atoi(argv[3]), H = atoi(argv[4]);
  const int frames = argc > 6 ? atoi(argv[6]) : 1;
  Renderer renderer(W, H);

  // Create scene (main.cpp:85-100)
  model_id_t box = renderer.load_model(argv[1]);
  model_id_t dynamic_object = renderer.load_model(argv[2]);
  if (box == (model_id_t)INVALID_ID || dynamic_object == (model_id_t)INVALID_ID) return 1;
  material_data_t dynamic_object_material;
  memset(&dynamic_object_material, 0, sizeof dynamic_object_material);
  dynamic_object_material.ambient = vct::vec4(1, 1, 1, 1);
  dynamic_object_material.diffuse = vct::vec4(0, 0, 0, 0);
  dynamic_object_material.specular = vct::vec4(1, 1, 1, 1);
  dynamic_object_material.transmittance = vct::vec4(1, 1, 1, 1);
  dynamic_object_material.emission = vct::vec3(0.0f, 0.0f, 0.25f);
  dynamic_object_material.shininess = 1000;
  dynamic_object_material.ior = 5;
  dynamic_object_material.dissolve = 0.1f;
  dynamic_object_material.illum = 4;
  material_id_t dynamic_object_material_id = renderer.add_material(dynamic_object_material);
  renderer.set_model_material(dynamic_object_material_id, dynamic_object);
  const V3 dynamic_object_position = {0.0f, 1.1f, -0.5f};
  const float dynamic_object_scale = 0.3f;

  renderer.set_grid_size(3);

  // Camera camera(vec3(0, .9, 3), 0, -90); set_perspective(45.0f, 4/3, 0.1, 100)
  const float pitch = 0.0f, yaw = -90.0f;
  const V3 front = normalize({cosf(radians(pitch)) * cosf(radians(yaw)), sinf(radians(pitch)),
                              cosf(radians(pitch)) * sinf(radians(yaw))});
  const V3 eye = {0.0f, 0.9f, 3.0f};
  vct::mat4 view = look_at(eye, {eye.x + front.x, eye.y + front.y, eye.z + front.z}, {0.0f, 1.0f, 0.0f});
  vct::mat4 projection = perspective(45.0f, 800.0f / 600.0f, 0.1f, 100.0f);
  renderer.set_camera_transform(view, projection);

  point_light_t ceiling_light;
  ceiling_light.position = vct::vec3(0.0f, 1.4f, 0.0f);
  ceiling_light.color = vct::vec3(1.0f, 1.0f, 1.0f);
  ceiling_light.intensity = 1.0f;

  for (int f = 0; f < frames; ++f) {  // the body of the frame loop (main.cpp:300-387)
    renderer.set_voxel_view_dir(7, 0.0f);
    renderer.upload_material_data(dynamic_object_material, dynamic_object_material_id);
    renderer.set_rendering_phases(true, true, true, true);
    vct::mat4 dynamic_object_matrix;  // translate * rotate(0) * scale
    dynamic_object_matrix.m[0] = dynamic_object_matrix.m[5] = dynamic_object_matrix.m[10] = dynamic_object_scale;
    dynamic_object_matrix.m[12] = dynamic_object_position.x;
    dynamic_object_matrix.m[13] = dynamic_object_position.y;
    dynamic_object_matrix.m[14] = dynamic_object_position.z;
    renderer.set_model_transform(dynamic_object, dynamic_object_matrix);
    vct::mat4 box_matrix;
    renderer.set_model_transform(box, box_matrix);
    renderer.queue_model(box);
    renderer.queue_model(dynamic_object);
    renderer.queue_point_light(ceiling_light);
    renderer.render();
  }
  // the drop-in's material table: main.cpp's material has its id as `ubo`
  const material_t& mt = renderer.get_material(dynamic_object_material_id);
  if (mt.ubo != (GLuint)dynamic_object_material_id || mt.num_textures != 0 || mt.data.illum != 4) {
    fprintf(stderr, "get_material returned an unexpected entry\n");
    return 1;
  }
  std::vector<float> img = renderer.image();
  FILE* fp = fopen(argv[5], "wb");
  if (!fp || fwrite(img.data(), sizeof(float), img.size(), fp) != img.size()) {
    fprintf(stderr, "cannot write %s\n", argv[5]);
    return 1;
  }
  fclose(fp);
  printf("rendered %d frame(s) %dx%d through Renderer; models %zu %zu, material %zu\n", frames, W, H, box,
         dynamic_object, dynamic_object_material_id);
  return 0;
}
