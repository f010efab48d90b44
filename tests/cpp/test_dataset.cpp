// The transforms.json loader and initial gaussians, without a GPU, restating
// the reference's loader test (engine/tests/test_trainer.cpp:909-1054): PNG
// frames written by hand (stored zlib blocks), intrinsics rescaled to the
// short side, holdout every 8th frame, opaque frames without masks, area
// resampling, camera centre and axes, random initialisation; a damaged frame
// dropped and a mostly unreadable dataset rejected; a capture with masks and
// partial depth + confidence initialising from depth.  Also: voxel merging
// over budget, Morton order, neighbour scales.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "gs/camera.h"
#include "gs/dataset.h"
#include "gs/image_io.h"
#include "test_util.h"

namespace fs = std::filesystem;

namespace {

void put(const fs::path& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  f << text;
}

// An 8-bit RGB PNG whose row y, byte x is value + x, in one stored (uncompressed) deflate block.
std::string stored_png(uint32_t w, uint32_t h, uint8_t value) {
  auto crc32 = [](const uint8_t* data, size_t n) {
    uint32_t crc = 0xffffffffu;
    for (size_t i = 0; i < n; ++i) {
      crc ^= data[i];
      for (int k = 0; k < 8; ++k) crc = (crc >> 1) ^ (0xedb88320u & (0u - (crc & 1u)));
    }
    return ~crc;
  };
  auto be32 = [](std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back(uint8_t(x >> s));
  };
  std::vector<uint8_t> raw;
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);
    for (uint32_t x = 0; x < w * 3; ++x) raw.push_back(uint8_t(value + x));
  }
  const uint16_t n = uint16_t(raw.size());
  std::vector<uint8_t> z = {0x78, 0x01, 1, uint8_t(n), uint8_t(n >> 8), uint8_t(~n), uint8_t(~n >> 8)};
  z.insert(z.end(), raw.begin(), raw.end());
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521;
    b = (b + a) % 65521;
  }
  be32(z, (b << 16) | a);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
    be32(out, uint32_t(data.size()));
    std::vector<uint8_t> body(type, type + 4);
    body.insert(body.end(), data.begin(), data.end());
    out.insert(out.end(), body.begin(), body.end());
    be32(out, crc32(body.data(), body.size()));
  };
  std::vector<uint8_t> ihdr;
  be32(ihdr, w);
  be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  chunk("IHDR", ihdr);
  chunk("IDAT", z);
  chunk("IEND", {});
  return std::string(out.begin(), out.end());
}

std::string bytes_of(const std::vector<uint8_t>& v) { return std::string(v.begin(), v.end()); }

void test_loader() {
  const fs::path dir = fs::temp_directory_path() / ("gs_dataset_test_" + std::to_string(::getpid()));
  fs::remove_all(dir);
  fs::create_directories(dir / "images");
  std::string frames;
  for (int i = 0; i < 9; ++i) {
    const std::string name = "images/f" + std::to_string(i) + ".png";
    put(dir / name, stored_png(8, 4, uint8_t(10 * i)));
    frames += std::string(i ? "," : "") + "{\"file_path\": \"" + name +
              "\", \"transform_matrix\": [[1,0,0," + std::to_string(i) + "],[0,1,0,0],[0,0,1,2],[0,0,0,1]]}";
  }
  put(dir / "transforms.json", "{\"w\": 8, \"h\": 4, \"fl_x\": 10.0, \"cx\": 4, \"cy\": 2, \"frames\": [" + frames + "]}");
  const gs::Dataset ds = gs::load_dataset(dir.string(), 2);
  EXPECT(ds.width == 4 && ds.height == 2, "dataset size %ux%u", ds.width, ds.height);
  EXPECT(std::fabs(ds.fx - 5.0f) < 1e-5f && std::fabs(ds.cx - 2.0f) < 1e-5f, "intrinsics not rescaled");
  EXPECT(ds.frames.size() == 9 && ds.holdout_count() == 2, "frames %zu, holdout %zu", ds.frames.size(),
         ds.holdout_count());
  EXPECT(!ds.has_masks && ds.frames[1].rgba[3] == 255, "frames without masks must be opaque");
  // Pixel (0, 0) of frame 1 averages source pixels (0..1, 0..1): red 10 and 13.
  EXPECT(std::abs(int(ds.frames[1].rgba[0]) - 12) <= 1, "resized red %d", ds.frames[1].rgba[0]);
  EXPECT(std::fabs(ds.camera_center[0] - 4.0f) < 1e-5f, "camera centre x %.3f", ds.camera_center[0]);
  const gs::Camera cam = gs::dataset_camera(ds, ds.frames[3]);
  EXPECT(std::fabs(cam.position[0] - 3.0f) < 1e-6f && std::fabs(cam.translation[0] + 3.0f) < 1e-6f,
         "camera translation wrong");
  EXPECT(cam.rotation[8] == -1.0f && cam.rotation[4] == -1.0f, "OpenGL -> OpenCV axes not flipped");
  gs::InitStats stats;
  gs::InitOptions init;
  init.max_gaussians = 100;
  const gs::SplatScene scene = gs::initialize_gaussians(ds, init, &stats);
  EXPECT(stats.source == "random" && scene.count == 100, "random init %s %u", stats.source.c_str(), scene.count);

  put(dir / "images/f4.png", "not a png");
  const gs::Dataset partial = gs::load_dataset(dir.string(), 2);
  EXPECT(partial.frames.size() == 8 && partial.skipped.size() == 1, "damaged frame: %zu frames, %zu skipped",
         partial.frames.size(), partial.skipped.size());
  EXPECT(partial.frames.size() > 4 && partial.frames[4].name == "images/f5.png", "wrong frame dropped");
  for (int i = 0; i < 6; ++i) put(dir / ("images/f" + std::to_string(i) + ".png"), "");
  bool threw = false;
  try {
    gs::load_dataset(dir.string(), 2);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  EXPECT(threw, "a dataset with most frames unreadable should be rejected");
  fs::remove_all(dir);

  // A capture: masks on every frame, depth (mm, half resolution) + confidence on every other.
  fs::create_directories(dir / "images");
  fs::create_directories(dir / "depth");
  frames.clear();
  for (int i = 0; i < 6; ++i) {
    const std::string k = std::to_string(i);
    const std::string name = "images/f" + k + ".png", mask = "depth/m" + k + ".png";
    put(dir / name, stored_png(8, 4, uint8_t(10 * i)));
    std::vector<uint8_t> m(8 * 4, 255);
    put(dir / mask, bytes_of(gs::encode_png_gray8(m.data(), 8, 4)));
    std::string entry = "{\"file_path\": \"" + name + "\", \"mask_path\": \"" + mask + "\"";
    if (i % 2 == 0) {
      std::vector<uint16_t> d(4 * 2, uint16_t(1500 + i));
      std::vector<uint8_t> c(4 * 2, 200);
      put(dir / ("depth/d" + k + ".png"), bytes_of(gs::encode_png_gray16(d.data(), 4, 2)));
      put(dir / ("depth/c" + k + ".png"), bytes_of(gs::encode_png_gray8(c.data(), 4, 2)));
      entry += ", \"depth_file_path\": \"depth/d" + k + ".png\", \"depth_confidence_path\": \"depth/c" + k + ".png\"";
    }
    frames += std::string(i ? "," : "") + entry + ", \"transform_matrix\": [[1,0,0," + std::to_string(0.1 * i) +
              "],[0,1,0,0],[0,0,1,2],[0,0,0,1]]}";
  }
  put(dir / "transforms.json", "{\"w\": 8, \"h\": 4, \"fl_x\": 10.0, \"cx\": 4, \"cy\": 2, "
                               "\"depth_unit_scale_factor\": 0.001, \"frames\": [" + frames + "]}");
  const gs::Dataset cap = gs::load_dataset(dir.string(), 2);
  size_t with_depth = 0;
  bool depth_ok = true;
  for (const gs::DatasetFrame& f : cap.frames) {
    if (f.depth.empty()) continue;
    ++with_depth;
    depth_ok = depth_ok && f.depth_width == 4 && f.depth_height == 2 && f.depth_confidence.size() == 8 &&
               f.depth_confidence[0] == 200 && std::fabs(f.depth[0] - 1.5f) < 0.01f;
  }
  EXPECT(cap.width == 4 && cap.height == 2, "capture size %ux%u", cap.width, cap.height);
  EXPECT(cap.has_depth && cap.has_masks && with_depth == 3 && depth_ok,
         "capture: depth %d, masks %d, %zu frames with depth, values ok %d", int(cap.has_depth), int(cap.has_masks),
         with_depth, int(depth_ok));
  gs::InitOptions ci;
  ci.max_gaussians = 1000;
  gs::InitStats cs;
  const gs::SplatScene from_depth = gs::initialize_gaussians(cap, ci, &cs);
  // Frames 2 and 4 train with depth (frame 0 is held out): 2 x 8 samples, 1.5 m in front of x = 0.2, 0.4.
  EXPECT(cs.source == "depth" && cs.raw_points == 16 && from_depth.count == 16, "capture init: %s, %zu raw, %u",
         cs.source.c_str(), cs.raw_points, from_depth.count);
  bool placed = true;
  for (uint32_t i = 0; i < from_depth.count; ++i) {
    // OpenCV z forward is OpenGL -z: the points sit at world z = 2 - depth.
    placed = placed && std::fabs(from_depth.positions[i * 3 + 2] - (2.0f - 1.502f)) < 0.01f;
  }
  EXPECT(placed, "depth points not 1.5 m in front of the cameras");
  fs::remove_all(dir);
}

void test_merging_and_scales() {
  // A 40 x 40 x 10 lattice of spacing 0.05 as a point cloud (the spacing is under 2% of its
  // extent, so scales are not clamped); budget 1000 forces voxel merging.
  gs::Dataset ds = gs::make_orbit_dataset(4, 16, 16, 3.0f, 0.0f, 60.0f);
  for (int x = 0; x < 40; ++x)
    for (int y = 0; y < 40; ++y)
      for (int z = 0; z < 10; ++z) {
        ds.points.insert(ds.points.end(), {0.05f * x, 0.05f * y, 0.05f * z});
        ds.point_colors.insert(ds.point_colors.end(), {uint8_t(5 * x), uint8_t(5 * y), uint8_t(20 * z)});
      }
  gs::InitOptions o;
  o.max_gaussians = 1000;
  gs::InitStats st;
  const gs::SplatScene s = gs::initialize_gaussians(ds, o, &st);
  EXPECT(st.source == "points" && st.raw_points == 16000 && st.voxel_size > 0.0f, "merge: %s %zu %.4f",
         st.source.c_str(), st.raw_points, st.voxel_size);
  EXPECT(s.count > 100 && s.count <= 1000, "merged count %u", s.count);
  // Under budget: every point kept, scale = neighbour spacing 0.05.
  o.max_gaussians = 20000;
  const gs::SplatScene k = gs::initialize_gaussians(ds, o, &st);
  EXPECT(k.count == 16000 && st.voxel_size == 0.0f, "kept %u", k.count);
  bool scales = true;
  for (uint32_t i = 0; i < k.count; ++i) scales = scales && std::fabs(std::exp(k.log_scales[i * 3]) - 0.05f) < 1e-3f;
  EXPECT(scales, "lattice scales should be the spacing");
  // Morton order: consecutive gaussians are close (mean step far below the cube's size).
  double step = 0.0;
  for (uint32_t i = 1; i < k.count; ++i) {
    double d2 = 0.0;
    for (int c = 0; c < 3; ++c) d2 += std::pow(k.positions[i * 3 + c] - k.positions[(i - 1) * 3 + c], 2);
    step += std::sqrt(d2);
  }
  EXPECT(step / (k.count - 1) < 0.2, "Morton order: mean step %.3f", step / (k.count - 1));
  EXPECT(std::fabs(k.opacity_logits[0] - std::log(0.2f / 0.8f)) < 1e-6f, "initial opacity");
}

}  // namespace

int main() {
  try {
    test_loader();
    test_merging_and_scales();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_dataset: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
