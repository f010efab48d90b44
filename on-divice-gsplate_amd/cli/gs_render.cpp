// Renders a splat file (or a synthetic scene) with the viewer's image contract
// (the command surface of engine/cli/gs_render.cpp:1-254).
//
//   gs_render <scene> <out.ppm> [width height yaw_deg pitch_deg] [--look x,y,z] [--distance d]
//       one orbit view of the scene (800x600, yaw 0, pitch 0.25 rad by default)
//   gs_render <scene> --views <cameras.txt> <out_dir>
//       renders every camera of cameras.txt to out_dir/view_NNNNN.ppm (N from
//       00000, in file order).  A camera is a line of 18 numbers: image width and
//       height, fx fy cx cy in pixels, and the world -> camera transform [R | t]
//       (camera axes x right, y down, z forward) as three rows of four; blank
//       lines are skipped.
//
// <scene> is a .ply/.spz file or synthetic:N.  The app's places and
// relighting (--place, --yaw, --relight) are not part of this build.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gs/camera.h"
#include "gs/device.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"

namespace {

constexpr float kClear[3] = {0.08f, 0.09f, 0.11f};

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// "synthetic[:N]" names the torus-knot demo scene of N splats (100000 when
// no count is given); anything else is a .ply or .spz file.
gs::SplatScene scene_from(const std::string& input) {
  const std::string tag = "synthetic";
  if (input.compare(0, tag.size(), tag) != 0) {
    const std::vector<uint8_t> bytes = gs::read_file(input);
    return gs::load_splat_file(bytes.data(), bytes.size());
  }
  const size_t colon = input.find(':');
  const uint32_t n = colon == std::string::npos ? 100000u : uint32_t(std::strtoul(input.c_str() + colon + 1, nullptr, 10));
  return gs::make_synthetic_scene(n);
}

// The scene and one line about it (the count, SH degree, axis convention,
// antialiasing flag and load time, as the reference's viewer reports them).
gs::SplatScene load_scene(const std::string& input) {
  const auto t0 = std::chrono::steady_clock::now();
  gs::SplatScene scene = scene_from(input);
  const char* axes = scene.convention == gs::AxisConvention::RDF ? "RDF" : "RUB";
  std::printf("loaded %u splats, sh degree %d, %s, antialiased %d (%.0f ms)\n", scene.count, scene.sh_degree, axes,
              int(scene.antialiased), ms_since(t0));
  return scene;
}

// Up to `max` leading numbers of `line` into out[]; returns how many numbers
// the line starts with, counting one past `max` when there are more.
int scan_numbers(const std::string& line, double* out, int max) {
  const char* p = line.c_str();
  int n = 0;
  for (char* end = nullptr; n <= max; p = end) {
    const double x = std::strtod(p, &end);
    if (end == p) break;
    if (n < max) out[n] = x;
    ++n;
  }
  return n;
}

// One cameras.txt row: width height fx fy cx cy, then [R | t] row by row.
gs::Camera camera_from_row(const double* v) {
  gs::Camera c{};
  c.width = uint32_t(v[0]);
  c.height = uint32_t(v[1]);
  c.fx = float(v[2]), c.fy = float(v[3]), c.cx = float(v[4]), c.cy = float(v[5]);
  for (int k = 0; k < 12; ++k) {
    const int row = k / 4, col = k % 4;
    (col == 3 ? c.translation[row] : c.rotation[row * 3 + col]) = float(v[6 + k]);
  }
  const float* R = c.rotation;
  const float* t = c.translation;
  for (int k = 0; k < 3; ++k) c.position[k] = -(R[k] * t[0] + R[3 + k] * t[1] + R[6 + k] * t[2]);  // -R^T t
  return c;
}

// The cameras of a cameras.txt file (the format of engine/cli/gs_render.cpp:145-175).
std::vector<gs::Camera> read_cameras(const std::string& path) {
  const std::vector<uint8_t> raw = gs::read_file(path);
  const std::string text(raw.begin(), raw.end());
  std::vector<gs::Camera> cameras;
  size_t start = 0;
  while (start < text.size()) {
    const size_t newline = std::min(text.find('\n', start), text.size());
    const std::string line = text.substr(start, newline - start);
    start = newline + 1;
    double row[18];
    const int n = scan_numbers(line, row, 18);
    if (n == 0) continue;
    if (n != 18) throw std::runtime_error("camera line needs 18 numbers: " + line);
    cameras.push_back(camera_from_row(row));
  }
  return cameras;
}

// Every camera to out_dir/view_NNNNN.ppm.  A file goes to the device as it is
// (SPZ is unpacked there), once per output size; synthetic scenes are built
// on the host.
void render_views(const std::string& input, const std::vector<gs::Camera>& cameras, const std::string& out_dir) {
  const bool synthetic = input.rfind("synthetic", 0) == 0;
  const gs::SplatScene scene = synthetic ? load_scene(input) : gs::SplatScene{};
  const std::vector<uint8_t> bytes = synthetic ? std::vector<uint8_t>{} : gs::read_file(input);
  gs::Device device(0);
  std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<gs::GpuScene>> by_size;
  gs::SplatRenderer renderer(device);
  const auto t0 = std::chrono::steady_clock::now();
  size_t done = 0;
  for (const gs::Camera& c : cameras) {
    std::unique_ptr<gs::GpuScene>& on_device = by_size[{c.width, c.height}];
    if (on_device == nullptr) {
      const auto t_up = std::chrono::steady_clock::now();
      on_device = synthetic ? gs::upload_scene(device, scene, c.width, c.height)
                            : gs::upload_splat_file(device, bytes.data(), bytes.size(), c.width, c.height);
      std::printf("scene on the device for %ux%u: %u splats, sh degree %d (%.0f ms)\n", c.width, c.height,
                  on_device->count(), on_device->sh_degree(), ms_since(t_up));
    }
    renderer.attach(on_device.get());
    const std::vector<uint8_t> pixels = renderer.render(c, kClear);
    char name[32];
    std::snprintf(name, sizeof(name), "/view_%05zu.ppm", done++);
    gs::write_ppm(out_dir + name, pixels, c.width, c.height);
    std::printf("\rview %zu/%zu (%u visible)", done, cameras.size(), renderer.visible_count());
    std::fflush(stdout);
  }
  std::printf("\n%zu views in %.1f ms\n", cameras.size(), ms_since(t0));
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
      render_views(input, read_cameras(argv[3]), argv[4]);
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
        // A point of the scene; the orbit turns Y-down data half a turn about X.
        float p[3] = {0, 0, 0};
        std::sscanf(argv[i + 1], "%f,%f,%f", &p[0], &p[1], &p[2]);
        const float turn = orbit.flip_y ? -1.0f : 1.0f;
        const float shown[3] = {p[0], turn * p[1], turn * p[2]};
        std::copy_n(shown, 3, orbit.target);
      }
      if (a == "--distance") orbit.distance = float(std::atof(argv[i + 1]));
    }
    auto gpu = gs::upload_scene(device, scene, width, height);
    gs::SplatRenderer renderer(device);
    renderer.attach(gpu.get());
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<uint8_t> pixels = renderer.render(orbit.camera(width, height), kClear);
    std::printf("rendered %ux%u: %u visible, %.1f ms on %s\n", width, height, renderer.visible_count(), ms_since(t0),
                device.name().c_str());
    gs::write_ppm(argv[2], pixels, width, height);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
