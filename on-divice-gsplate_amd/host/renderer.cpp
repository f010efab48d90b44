// Scene upload and rendering for gs_render (the role of
// engine/src/splat_renderer.cpp:67-348 and offscreen.cpp): the scene goes
// into a workspace once, its SH rest rounded to fp16 as the viewer stores it
// (splat_renderer.cpp:103-116); each render is the viewer's image contract
// (quads back to front into an 8-bit target over the clear colour) through
// gs_hip_view_render.
#include "gs/renderer.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "gs_hip.h"

namespace gs {

float round_to_half(float f) {
  if (!std::isfinite(f)) return f;
  const float a = std::fabs(f);
  if (a >= 65520.0f) return std::copysign(INFINITY, f);
  int e;
  std::frexp(a, &e);                         // a = m 2^e, m in [0.5, 1)
  const int step = std::max(e - 11, -24);   // 11 significant bits; subnormal spacing 2^-24
  const double q = std::nearbyint(std::ldexp(double(a), -step));  // default rounding: ties to even
  return std::copysign(float(std::ldexp(q, step)), f);
}

GpuScene::GpuScene(Device& device, const SplatScene& scene, uint32_t width, uint32_t height)
    : count_(scene.count), width_(width), height_(height), sh_degree_(scene.sh_degree),
      antialiased_(scene.antialiased) {
  (void)device;
  check_status(gs_hip_workspace_create(&ws_, std::max<uint32_t>(scene.count, 1), scene.sh_degree, width, height));
  std::vector<float> p(size_t(scene.count) * 16, 0.0f);
  for (uint32_t i = 0; i < scene.count; ++i) {
    float* d = &p[size_t(i) * 16];
    for (int k = 0; k < 3; ++k) {
      d[k] = scene.positions[i * 3 + k];
      d[4 + k] = scene.log_scales[i * 3 + k];
      d[12 + k] = scene.sh_dc[i * 3 + k];
    }
    d[3] = scene.opacity_logits[i];
    d[7] = 1.0f;
    for (int k = 0; k < 4; ++k) d[8 + k] = scene.rotations[i * 4 + k];
  }
  const int coeffs = sh_rest_coeffs(scene.sh_degree);
  std::vector<float> rest(scene.sh_rest.begin(), scene.sh_rest.end());
  for (float& r : rest) r = round_to_half(r);
  const int rc = gs_hip_workspace_load(ws_, p.data(), coeffs > 0 ? rest.data() : nullptr, scene.count,
                                       uint32_t(coeffs));
  if (rc != 0) {
    char msg[512];
    gs_hip_last_error(msg, sizeof(msg));
    gs_hip_workspace_destroy(ws_);
    throw std::runtime_error(msg);
  }
}

GpuScene::GpuScene(Device& device, const uint8_t* spz, size_t size, uint32_t width, uint32_t height)
    : width_(width), height_(height) {
  (void)device;
  int degree = 0, aa = 0;
  check_status(gs_hip_spz_info(spz, size, &count_, &degree, &aa));
  check_status(gs_hip_workspace_create(&ws_, std::max<uint32_t>(count_, 1), degree, width, height));
  // SPZ's SH rest values are k / 128 with |k| <= 128: already fp16 values, as the viewer stores them.
  const int rc = gs_hip_workspace_load_spz(ws_, spz, size, &count_, &sh_degree_, &aa);
  if (rc != 0) {
    char msg[512];
    gs_hip_last_error(msg, sizeof(msg));
    gs_hip_workspace_destroy(ws_);
    throw std::runtime_error(msg);
  }
  antialiased_ = aa != 0;
}

GpuScene::~GpuScene() {
  if (ws_ != nullptr) gs_hip_workspace_destroy(ws_);
}

std::unique_ptr<GpuScene> upload_scene(Device& device, const SplatScene& scene, uint32_t width, uint32_t height) {
  return std::make_unique<GpuScene>(device, scene, width, height);
}

std::unique_ptr<GpuScene> upload_splat_file(Device& device, const uint8_t* data, size_t size, uint32_t width,
                                            uint32_t height) {
  if (size >= 2 && data[0] == 0x1f && data[1] == 0x8b) return std::make_unique<GpuScene>(device, data, size, width, height);
  return std::make_unique<GpuScene>(device, load_splat_file(data, size), width, height);
}

std::vector<uint8_t> SplatRenderer::render(const Camera& camera, const float clear[3]) {
  if (scene_ == nullptr) throw std::logic_error("SplatRenderer::render: no scene attached");
  if (camera.width != scene_->width() || camera.height != scene_->height()) {
    throw std::invalid_argument("SplatRenderer::render: camera size differs from the scene's output size");
  }
  gs_hip_frame f;
  std::memset(&f, 0, sizeof(f));
  std::memcpy(f.cam.rotation, camera.rotation, sizeof(f.cam.rotation));
  std::memcpy(f.cam.translation, camera.translation, sizeof(f.cam.translation));
  std::memcpy(f.cam.position, camera.position, sizeof(f.cam.position));
  f.cam.fx = camera.fx, f.cam.fy = camera.fy, f.cam.cx = camera.cx, f.cam.cy = camera.cy;
  f.cam.width = camera.width, f.cam.height = camera.height;
  f.tiles_x = (camera.width + GS_HIP_TILE - 1) / GS_HIP_TILE;
  f.tiles_y = (camera.height + GS_HIP_TILE - 1) / GS_HIP_TILE;
  f.slots = scene_->count();
  f.sh_degree = uint32_t(scene_->sh_degree());
  const float rgba[4] = {clear[0], clear[1], clear[2], 1.0f};
  std::vector<uint8_t> out(size_t(camera.width) * camera.height * 4);
  check_status(gs_hip_view_render(scene_->workspace(), &f, rgba, scene_->antialiased() ? GS_HIP_VIEW_ANTIALIASED : 0u,
                                  out.data(), &visible_));
  return out;
}

std::vector<uint8_t> encode_ppm(const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height) {
  if (rgba.size() != size_t(width) * height * 4) throw std::invalid_argument("encode_ppm: size mismatch");
  const std::string header = "P6\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
  std::vector<uint8_t> out(header.begin(), header.end());
  out.reserve(out.size() + size_t(width) * height * 3);
  for (size_t p = 0; p < size_t(width) * height; ++p) out.insert(out.end(), &rgba[p * 4], &rgba[p * 4 + 3]);
  return out;
}

void write_ppm(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height) {
  write_file(path, encode_ppm(rgba, width, height));
}

}  // namespace gs
