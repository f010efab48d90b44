// Rendering a splat scene to an RGBA8 image on the GPU (the gs_render path:
// engine/include/gs/splat_renderer.h:63-114 and offscreen.h:14-44), with
// the Vulkan render-pass types replaced by a plain render() call.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "gs/camera.h"
#include "gs/device.h"
#include "gs/splat_scene.h"

struct gs_hip_workspace;

namespace gs {

// A scene resident on the device at a fixed output size.
class GpuScene {
 public:
  GpuScene(Device& device, const SplatScene& scene, uint32_t width, uint32_t height);
  // An SPZ file decoded on the device (gs_hip_workspace_load_spz): the same
  // scene as GpuScene(device, load_spz(data, size), ...), without the host
  // unpacking.
  GpuScene(Device& device, const uint8_t* spz, size_t size, uint32_t width, uint32_t height);
  ~GpuScene();
  GpuScene(const GpuScene&) = delete;
  GpuScene& operator=(const GpuScene&) = delete;
  uint32_t count() const { return count_; }
  int sh_degree() const { return sh_degree_; }
  bool antialiased() const { return antialiased_; }
  uint32_t width() const { return width_; }
  uint32_t height() const { return height_; }
  gs_hip_workspace* workspace() { return ws_; }

 private:
  gs_hip_workspace* ws_ = nullptr;
  uint32_t count_ = 0, width_ = 0, height_ = 0;
  int sh_degree_ = 0;
  bool antialiased_ = false;
};

std::unique_ptr<GpuScene> upload_scene(Device& device, const SplatScene& scene, uint32_t width, uint32_t height);
// A .ply or .spz file's scene on the device: SPZ decoded there, PLY through load_ply.
std::unique_ptr<GpuScene> upload_splat_file(Device& device, const uint8_t* data, size_t size, uint32_t width,
                                            uint32_t height);

class SplatRenderer {
 public:
  explicit SplatRenderer(Device& device) : device_(device) {}
  void attach(GpuScene* scene) { scene_ = scene; }
  // RGBA8 of `camera` (the scene's output size): splats back to front over
  // the opaque clear colour; visible_count() = splats drawn.
  std::vector<uint8_t> render(const Camera& camera, const float clear[3]);
  uint32_t visible_count() const { return visible_; }

 private:
  [[maybe_unused]] Device& device_;
  GpuScene* scene_ = nullptr;
  uint32_t visible_ = 0;
};

// The nearest fp16 value to `f` (ties to even) as a float: the precision at
// which the viewer stores SH rest coefficients (splat_renderer.cpp:103-116).
float round_to_half(float f);

// Binary PPM (P6) of an RGBA8 image.
std::vector<uint8_t> encode_ppm(const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height);
void write_ppm(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height);

}  // namespace gs
