// Renders a splat file (or a synthetic scene) with the viewer's image contract
// (the command surface of engine/cli/gs_render.cpp:1-254).
//
//   gs_render <scene> <out.ppm> [width height yaw_deg pitch_deg] [--look x,y,z] [--distance d]
//       one orbit view of the scene (800x600, yaw 0, pitch 0.25 rad by default)
//   gs_render <scene> --views <cameras.txt> <out_dir>
//       one image per line of cameras.txt, written as out_dir/view_00000.ppm, ...
//       Each line: width height fx fy cx cy, then the 3x4 world-to-camera matrix
//       (OpenCV convention: x right, y down, z forward) row by row.
//
// <scene> is a .ply/.spz file or synthetic:N.  The app's places and
// relighting (--place, --yaw, --relight) are not part of this build.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "gs/camera.h"
#include "gs/device.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"

namespace {

constexpr float kClear[3] = {0.08f, 0.09f, 0.11f};

gs::SplatScene load_scene(const std::string& input) {
  const auto t0 = std::chrono::steady_clock::now();
  gs::SplatScene scene;
  if (input.rfind("synthetic", 0) == 0) {
    const uint32_t n = input.size() > 10 ? uint32_t(std::atoi(input.c_str() + 10)) : 100000u;
    scene = gs::make_synthetic_scene(n);
  } else {
    const std::vector<uint8_t> bytes = gs::read_file(input);
    scene = gs::load_splat_file(bytes.data(), bytes.size());
  }
  std::printf("loaded %u splats, sh degree %d, %s, antialiased %d (%.0f ms)\n", scene.count, scene.sh_degree,
              scene.convention == gs::AxisConvention::RDF ? "RDF" : "RUB", scene.antialiased ? 1 : 0,
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  return scene;
}

std::vector<gs::Camera> read_cameras(const std::string& path) {
  std::ifstream file(path);
  if (!file) throw std::runtime_error("cannot open " + path);
  std::vector<gs::Camera> cameras;
  std::string line;
  while (std::getline(file, line)) {
    std::istringstream in(line);
    std::vector<double> v;
    double x;
    while (v.size() < 19 && in >> x) v.push_back(x);
    if (v.empty()) continue;
    if (v.size() != 18) throw std::runtime_error("camera line needs 18 numbers: " + line);
    gs::Camera c{};
    c.width = uint32_t(v[0]);
    c.height = uint32_t(v[1]);
    c.fx = float(v[2]), c.fy = float(v[3]), c.cx = float(v[4]), c.cy = float(v[5]);
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) c.rotation[r * 3 + k] = float(v[6 + r * 4 + k]);
      c.translation[r] = float(v[6 + r * 4 + 3]);
    }
    for (int k = 0; k < 3; ++k) {  // centre = -R^T t
      c.position[k] = -(c.rotation[k] * c.translation[0] + c.rotation[3 + k] * c.translation[1] +
                        c.rotation[6 + k] * c.translation[2]);
    }
    cameras.push_back(c);
  }
  return cameras;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr,
                 "usage: %s <scene.ply|scene.spz|synthetic:N> <out.ppm> [width height yaw_deg pitch_deg] "
                 "[--look x,y,z] [--distance d]\n"
                 "       %s <scene> --views <cameras.txt> <out_dir>\n",
                 argv[0], argv[0]);
    return 2;
  }
  try {
    const std::string mode = argv[2];
    const std::string input = argv[1];
    if (mode == "--views") {
      if (argc < 5) throw std::runtime_error("--views needs <cameras.txt> <out_dir>");
      const std::vector<gs::Camera> cameras = read_cameras(argv[3]);
      const std::string out_dir = argv[4];
      // A file goes to the device as it is (SPZ is unpacked there); synthetic scenes are built on the host.
      const bool synthetic = input.rfind("synthetic", 0) == 0;
      gs::SplatScene scene;
      std::vector<uint8_t> bytes;
      if (synthetic) {
        scene = load_scene(input);
      } else {
        bytes = gs::read_file(input);
      }
      gs::Device device(0);
      // One device copy of the scene per output size.
      std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<gs::GpuScene>> scenes;
      gs::SplatRenderer renderer(device);
      const auto t0 = std::chrono::steady_clock::now();
      for (size_t i = 0; i < cameras.size(); ++i) {
        const gs::Camera& c = cameras[i];
        auto& gpu = scenes[{c.width, c.height}];
        if (!gpu) {
          const auto tu = std::chrono::steady_clock::now();
          gpu = synthetic ? gs::upload_scene(device, scene, c.width, c.height)
                          : gs::upload_splat_file(device, bytes.data(), bytes.size(), c.width, c.height);
          std::printf("scene on the device for %ux%u: %u splats, sh degree %d (%.0f ms)\n", c.width, c.height,
                      gpu->count(), gpu->sh_degree(),
                      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tu).count());
        }
        renderer.attach(gpu.get());
        const std::vector<uint8_t> pixels = renderer.render(c, kClear);
        char name[32];
        std::snprintf(name, sizeof(name), "/view_%05zu.ppm", i);
        gs::write_ppm(out_dir + name, pixels, c.width, c.height);
        std::printf("\rview %zu/%zu (%u visible)", i + 1, cameras.size(), renderer.visible_count());
        std::fflush(stdout);
      }
      std::printf("\n%zu views in %.1f ms\n", cameras.size(),
                  std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
      return 0;
    }
    gs::SplatScene scene = load_scene(input);
    gs::Device device(0);
    auto positional = [&](int i) {
      return argc > i && (argv[i][0] != '-' || std::isdigit(static_cast<unsigned char>(argv[i][1])));
    };
    const uint32_t width = positional(3) ? uint32_t(std::atoi(argv[3])) : 800u;
    const uint32_t height = positional(4) ? uint32_t(std::atoi(argv[4])) : 600u;
    gs::OrbitCamera orbit;
    orbit.reset(gs::compute_bounds(scene), scene.convention == gs::AxisConvention::RDF);
    orbit.yaw = positional(5) ? float(std::atof(argv[5])) * 3.14159265f / 180.0f : 0.0f;
    orbit.pitch = positional(6) ? float(std::atof(argv[6])) * 3.14159265f / 180.0f : 0.25f;
    for (int i = 3; i + 1 < argc; ++i) {  // close-ups
      const std::string a = argv[i];
      if (a == "--look") {
        float q[3] = {0, 0, 0};
        std::sscanf(argv[i + 1], "%f,%f,%f", &q[0], &q[1], &q[2]);
        const float s = orbit.flip_y ? -1.0f : 1.0f;
        orbit.target[0] = q[0];
        orbit.target[1] = s * q[1];
        orbit.target[2] = s * q[2];
      }
      if (a == "--distance") orbit.distance = float(std::atof(argv[i + 1]));
    }
    auto gpu = gs::upload_scene(device, scene, width, height);
    gs::SplatRenderer renderer(device);
    renderer.attach(gpu.get());
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<uint8_t> pixels = renderer.render(orbit.camera(width, height), kClear);
    std::printf("rendered %ux%u: %u visible, %.1f ms on %s\n", width, height, renderer.visible_count(),
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                device.name().c_str());
    gs::write_ppm(argv[2], pixels, width, height);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
