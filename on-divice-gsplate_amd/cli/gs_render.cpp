// gs_render <scene.ply|scene.spz|synthetic[:N]> <out.ppm> [width height yaw_deg pitch_deg]
// Renders a splat scene from an orbit camera around its bounds into a PPM
// (the command surface of engine/cli/gs_render.cpp).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "gs/camera.h"
#include "gs/device.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <scene.ply|scene.spz|synthetic[:N]> <out.ppm> [width height yaw pitch]\n",
                 argv[0]);
    return 2;
  }
  try {
    const std::string in = argv[1], out = argv[2];
    const uint32_t width = argc > 3 ? uint32_t(std::atoi(argv[3])) : 1280;
    const uint32_t height = argc > 4 ? uint32_t(std::atoi(argv[4])) : 720;
    const float yaw = argc > 5 ? float(std::atof(argv[5])) : 0.0f;
    const float pitch = argc > 6 ? float(std::atof(argv[6])) : 15.0f;
    gs::SplatScene scene;
    if (in.rfind("synthetic", 0) == 0) {
      const uint32_t n = in.size() > 10 ? uint32_t(std::atoi(in.c_str() + 10)) : 100000;
      scene = gs::make_synthetic_scene(n, 1);
    } else {
      const auto bytes = gs::read_file(in);
      scene = gs::load_splat_file(bytes.data(), bytes.size());
    }
    // Y-down data (PLY, RDF) is shown upright by the orbit's flip.
    gs::OrbitCamera orbit;
    orbit.reset(gs::compute_bounds(scene), scene.convention == gs::AxisConvention::RDF);
    orbit.yaw = yaw * 3.14159265f / 180.0f;
    orbit.pitch = pitch * 3.14159265f / 180.0f;
    gs::Device device(0);
    auto gpu = gs::upload_scene(device, scene, width, height);
    gs::SplatRenderer renderer(device);
    renderer.attach(gpu.get());
    const float clear[3] = {0.08f, 0.09f, 0.11f};
    const auto rgba = renderer.render(orbit.camera(width, height), clear);
    gs::write_ppm(out, rgba, width, height);
    std::printf("%s: %u splats (%u visible) -> %s (%ux%u) on %s\n", in.c_str(), scene.count, renderer.visible_count(),
                out.c_str(), width, height, device.name().c_str());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gs_render: %s\n", e.what());
    return 1;
  }
}
