// vct_demo.cpp - headless driver of the Renderer API (the frame loop of
// src/main.cpp:141-388 without window, UI and input).
//
//   vct_demo [--frames N] [--res R] [--size WxH] [--obj file.obj] [--ppm out.ppm]
//            [--raw out.f32]
//
// Scene: the procedural Cornell box (or an OBJ file through load_model) and a
// ceiling point light, as main.cpp:85-121 sets up its scene; camera from a
// look-at, as Camera::get_lookat does.  Prints per-pass times and a checksum
// of the final image.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/vct_renderer.h"

static vct::mat4 look_at(const float eye[3], const float c[3], const float up[3]) {
  float f[3] = {c[0] - eye[0], c[1] - eye[1], c[2] - eye[2]};
  float fl = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (float& v : f) v /= fl;
  float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (float& v : s) v /= sl;
  float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  vct::mat4 m;
  float* M = m.m;
  M[0] = s[0]; M[4] = s[1]; M[8] = s[2];
  M[1] = u[0]; M[5] = u[1]; M[9] = u[2];
  M[2] = -f[0]; M[6] = -f[1]; M[10] = -f[2];
  M[3] = M[7] = M[11] = 0;
  M[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
  M[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
  M[14] = f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2];
  M[15] = 1;
  return m;
}

static vct::mat4 perspective(float fovy, float aspect, float n, float f) {
  vct::mat4 m;
  float* M = m.m;
  for (int i = 0; i < 16; ++i) M[i] = 0;
  const float t = tanf(fovy / 2);
  M[0] = 1 / (aspect * t);
  M[5] = 1 / t;
  M[10] = -(f + n) / (f - n);
  M[11] = -1;
  M[14] = -(2 * f * n) / (f - n);
  return m;
}

int main(int argc, char** argv) {
  int frames = 3, res = 64, W = 512, H = 512;
  std::string obj, ppm, raw;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--frames" && i + 1 < argc) frames = atoi(argv[++i]);
    else if (a == "--res" && i + 1 < argc) res = atoi(argv[++i]);
    else if (a == "--size" && i + 1 < argc) sscanf(argv[++i], "%dx%d", &W, &H);
    else if (a == "--obj" && i + 1 < argc) obj = argv[++i];
    else if (a == "--ppm" && i + 1 < argc) ppm = argv[++i];
    else if (a == "--raw" && i + 1 < argc) raw = argv[++i];
  }
  Renderer renderer(W, H);
  renderer.set_grid_resolution((unsigned)res);
  renderer.set_grid_size(0.55f);
  model_id_t box;
  if (!obj.empty()) {
    box = renderer.load_model(obj.c_str());
    if (box == (model_id_t)INVALID_ID) return 1;
  } else {
    std::vector<float> pos(34 * 9), nrm(34 * 9);
    std::vector<uint32_t> tm(34);
    vct_material mats[5];
    if (vct_scene_cornell(1, pos.data(), nrm.data(), tm.data(), mats) != VCT_OK) return 1;
    material_id_t base = 0;
    for (int i = 0; i < 5; ++i) {
      material_data_t d;
      memcpy(&d, &mats[i], sizeof d);
      material_id_t id = renderer.add_material(d);
      if (i == 0) base = id;
    }
    for (uint32_t& t : tm) t += (uint32_t)base;
    box = renderer.add_model(pos.data(), nrm.data(), tm.data(), 34);
  }
  const float eye[3] = {0.0f, 0.0f, 1.55f}, center[3] = {0, 0, 0}, up[3] = {0, 1, 0};
  vct::mat4 view = look_at(eye, center, up);
  vct::mat4 proj = perspective(45.0f * 3.14159265f / 180.0f, (float)W / (float)H, 0.1f, 100.0f);
  renderer.set_camera_transform(view, proj);
  renderer.set_rendering_phases(true, true, true, true);
  renderer.set_voxel_view_dir(7, 0.0f);  // main.cpp:300
  point_light_t light;
  light.position = vct::vec3(0.0f, 0.45f, 0.0f);
  light.color = vct::vec3(1.0f, 1.0f, 1.0f);
  light.intensity = 1.0f;
  vct::mat4 identity;
  double total = 0;
  for (int f = 0; f < frames; ++f) {
    renderer.set_model_transform(box, identity);
    renderer.queue_model(box);
    renderer.queue_point_light(light);
    auto t0 = std::chrono::steady_clock::now();
    renderer.render();
    vct_sync(renderer.context());
    auto t1 = std::chrono::steady_clock::now();
    total += std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  float ms[5];
  vct_get_pass_times(renderer.context(), ms, 5);
  std::vector<float> img = renderer.image();
  double sum = 0;
  for (size_t i = 0; i < img.size(); i += 4) sum += img[i] + img[i + 1] + img[i + 2];
  printf("frames %d avg %.3f ms  voxelize %.3f mip %.3f gbuffer %.3f trace %.3f ms  checksum %.6f\n", frames,
         total / frames, ms[0], ms[1], ms[2], ms[3], sum);
  if (!raw.empty()) {
    FILE* fp = fopen(raw.c_str(), "wb");
    if (fp) {
      fwrite(img.data(), 4, img.size(), fp);
      fclose(fp);
    }
  }
  if (!ppm.empty()) {
    FILE* fp = fopen(ppm.c_str(), "wb");
    if (fp) {
      fprintf(fp, "P6\n%d %d\n255\n", W, H);
      for (int y = H - 1; y >= 0; --y)
        for (int x = 0; x < W; ++x)
          for (int c = 0; c < 3; ++c) {
            float v = img[((size_t)y * W + x) * 4 + c];
            v = v < 0 ? 0 : (v > 1 ? 1 : v);
            fputc((int)(powf(v, 1.0f / 2.2f) * 255.0f + 0.5f), fp);
          }
      fclose(fp);
    }
  }
  return 0;
}
