// The dataset loader and the initial gaussians, without a GPU.  The criteria
// are the reference's loader test's (engine/tests/test_trainer.cpp:909-1054):
// intrinsics follow the resize to the short side, every 8th frame is held
// out, frames without masks are opaque, images are area-resampled, cameras
// come out in OpenCV axes around their mean, a frame that cannot be read is
// dropped while a mostly unreadable dataset is refused, and a capture with
// masks and partial depth + confidence initialises from its depth.  Also:
// voxel merging over budget, Morton order and neighbour scales.  The
// datasets are written here with the library's own PNG encoder.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "gs/camera.h"
#include "gs/dataset.h"
#include "gs/image_io.h"
#include "test_util.h"

namespace fs = std::filesystem;

namespace {

// A scratch dataset folder, removed when the test is done with it.
class Folder {
 public:
  explicit Folder(const std::string& tag)
      : root_(fs::temp_directory_path() / ("gs_ds_" + tag + "_" + std::to_string(::getpid()))) {
    fs::remove_all(root_);
    fs::create_directories(root_ / "rgb");
    fs::create_directories(root_ / "aux");
  }
  ~Folder() { fs::remove_all(root_); }
  std::string path() const { return root_.string(); }
  void write(const std::string& rel, const std::vector<uint8_t>& bytes) const {
    std::ofstream(root_ / rel, std::ios::binary).write(reinterpret_cast<const char*>(bytes.data()), std::streamsize(bytes.size()));
  }
  void write(const std::string& rel, const std::string& text) const { write(rel, std::vector<uint8_t>(text.begin(), text.end())); }

 private:
  fs::path root_;
};

// Source image k: 8 x 4 RGBA whose red at (x, y) is 10 k + 3 x + 7 y (green
// and blue other ramps), opaque.
constexpr uint32_t kW = 8, kH = 4;
uint8_t red_at(int k, uint32_t x, uint32_t y) { return uint8_t(10 * k + 3 * x + 7 * y); }

std::vector<uint8_t> source_png(int k) {
  std::vector<uint8_t> px(kW * kH * 4);
  for (uint32_t y = 0; y < kH; ++y)
    for (uint32_t x = 0; x < kW; ++x) {
      uint8_t* p = &px[(y * kW + x) * 4];
      p[0] = red_at(k, x, y), p[1] = uint8_t(100 + x), p[2] = uint8_t(200 - y), p[3] = 255;
    }
  return gs::encode_png_rgba8(px.data(), kW, kH);
}

// transforms.json with the given frame entries (already JSON objects).
std::string manifest(const std::vector<std::string>& frames, const std::string& extra = "") {
  std::ostringstream s;
  s << "{\"w\": " << kW << ", \"h\": " << kH << ", \"fl_x\": 10.0, \"cx\": 4.0, \"cy\": 2.0" << extra
    << ", \"frames\": [";
  for (size_t i = 0; i < frames.size(); ++i) s << (i ? ", " : "") << frames[i];
  s << "]}";
  return s.str();
}

// A camera at (tx, 0, 2) looking down -z (identity rotation, OpenGL axes).
std::string frame_entry(const std::string& image, double tx, const std::string& extra = "") {
  std::ostringstream s;
  s << "{\"file_path\": \"" << image << "\"" << extra << ", \"transform_matrix\": [[1, 0, 0, " << tx
    << "], [0, 1, 0, 0], [0, 0, 1, 2], [0, 0, 0, 1]]}";
  return s.str();
}

bool throws_runtime_error(const std::string& dir) {
  try {
    gs::load_dataset(dir, 2);
  } catch (const std::runtime_error&) {
    return true;
  }
  return false;
}

void test_plain_frames() {
  Folder f("plain");
  std::vector<std::string> frames;
  for (int k = 0; k < 9; ++k) {
    const std::string name = "rgb/" + std::to_string(k) + ".png";
    f.write(name, source_png(k));
    frames.push_back(frame_entry(name, double(k)));
  }
  f.write("transforms.json", manifest(frames));
  const gs::Dataset ds = gs::load_dataset(f.path(), 2);
  // 8 x 4 with the short side brought to 2: 4 x 2, focal and centre halved.
  EXPECT(ds.width == 4 && ds.height == 2, "training size %ux%u, want 4x2", ds.width, ds.height);
  EXPECT(std::fabs(ds.fx - 5.0f) < 1e-6f && std::fabs(ds.fy - 5.0f) < 1e-6f && std::fabs(ds.cx - 2.0f) < 1e-6f &&
             std::fabs(ds.cy - 1.0f) < 1e-6f,
         "intrinsics %g %g %g %g, want 5 5 2 1", ds.fx, ds.fy, ds.cx, ds.cy);
  size_t held = 0;
  for (size_t i = 0; i < ds.frames.size(); ++i) held += ds.frames[i].holdout == (i % 8 == 0) && ds.frames[i].holdout;
  EXPECT(ds.frames.size() == 9 && held == 2 && ds.holdout_count() == 2, "%zu frames, frames 0 and 8 held out: %zu",
         ds.frames.size(), held);
  bool opaque = !ds.has_masks;
  for (const gs::DatasetFrame& fr : ds.frames)
    for (size_t p = 0; p < 8; ++p) opaque = opaque && fr.rgba[p * 4 + 3] == 255;
  EXPECT(opaque, "frames without masks must be opaque");
  // Each training pixel is the mean of a 2 x 2 source block.
  bool means = true;
  for (int k : {1, 5})
    for (uint32_t y = 0; y < 2; ++y)
      for (uint32_t x = 0; x < 4; ++x) {
        const double want = (red_at(k, 2 * x, 2 * y) + red_at(k, 2 * x + 1, 2 * y) + red_at(k, 2 * x, 2 * y + 1) +
                             red_at(k, 2 * x + 1, 2 * y + 1)) / 4.0;
        means = means && std::fabs(ds.frames[size_t(k)].rgba[(y * 4 + x) * 4] - want) <= 1.0;
      }
  EXPECT(means, "area resampling: red is not the 2x2 block mean");
  // Cameras: OpenCV axes (y and z flipped), centred on x = 4 (the mean of 0..8).
  EXPECT(std::fabs(ds.camera_center[0] - 4.0f) < 1e-5f && std::fabs(ds.camera_center[2] - 2.0f) < 1e-5f,
         "camera centre (%g, %g, %g)", ds.camera_center[0], ds.camera_center[1], ds.camera_center[2]);
  for (size_t k : {size_t(0), size_t(3), size_t(7)}) {
    const gs::Camera c = gs::dataset_camera(ds, ds.frames[k]);
    EXPECT(c.rotation[0] == 1.0f && c.rotation[4] == -1.0f && c.rotation[8] == -1.0f, "frame %zu: axes", k);
    EXPECT(std::fabs(c.position[0] - float(k)) < 1e-6f && std::fabs(c.translation[0] + float(k)) < 1e-6f &&
               std::fabs(c.translation[2] - 2.0f) < 1e-6f,
           "frame %zu: centre %g, translation %g %g", k, c.position[0], c.translation[0], c.translation[2]);
  }
  gs::InitOptions init;
  init.max_gaussians = 100;
  gs::InitStats st;
  const gs::SplatScene scene = gs::initialize_gaussians(ds, init, &st);
  EXPECT(st.source == "random" && scene.count == 100 && st.raw_points == 100, "random init: %s, %u gaussians",
         st.source.c_str(), scene.count);
  bool inside = true;
  for (uint32_t i = 0; i < scene.count; ++i) {
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) d2 += std::pow(scene.positions[i * 3 + k] - ds.camera_center[k], 2);
    inside = inside && std::sqrt(d2) <= 0.5 * ds.camera_extent + 1e-4;
  }
  EXPECT(inside, "random points must lie within half the camera extent of the cameras' centre");

  // One unreadable image: that frame is dropped (and named), the rest keep their order.
  f.write("rgb/4.png", std::string("this is not an image"));
  const gs::Dataset partial = gs::load_dataset(f.path(), 2);
  EXPECT(partial.frames.size() == 8 && partial.skipped.size() == 1 &&
             partial.skipped[0].rfind("rgb/4.png", 0) == 0 && partial.frames[4].name == "rgb/5.png",
         "damaged frame: %zu frames, %zu skipped", partial.frames.size(), partial.skipped.size());
  // Five of nine unreadable: refused.
  for (int k : {0, 1, 2, 3}) f.write("rgb/" + std::to_string(k) + ".png", std::string());
  EXPECT(throws_runtime_error(f.path()), "a dataset with most frames unreadable must be refused");
}

void test_capture_with_depth() {
  Folder f("capture");
  std::vector<std::string> frames;
  const std::vector<uint8_t> mask(kW * kH, 255);
  for (int k = 0; k < 6; ++k) {
    const std::string id = std::to_string(k), image = "rgb/" + id + ".png", m = "aux/mask" + id + ".png";
    f.write(image, source_png(k));
    f.write(m, gs::encode_png_gray8(mask.data(), kW, kH));
    std::string extra = ", \"mask_path\": \"" + m + "\"";
    if (k % 2 == 0) {  // depth at half resolution, in millimetres, with a confidence map
      const std::vector<uint16_t> depth(4 * 2, uint16_t(1500 + k));
      const std::vector<uint8_t> conf(4 * 2, 200);
      f.write("aux/depth" + id + ".png", gs::encode_png_gray16(depth.data(), 4, 2));
      f.write("aux/conf" + id + ".png", gs::encode_png_gray8(conf.data(), 4, 2));
      extra += ", \"depth_file_path\": \"aux/depth" + id + ".png\", \"depth_confidence_path\": \"aux/conf" + id + ".png\"";
    }
    frames.push_back(frame_entry(image, 0.1 * k, extra));
  }
  f.write("transforms.json", manifest(frames, ", \"depth_unit_scale_factor\": 0.001"));
  const gs::Dataset ds = gs::load_dataset(f.path(), 2);
  size_t depth_frames = 0;
  bool samples = true;
  for (const gs::DatasetFrame& fr : ds.frames) {
    if (fr.depth.empty()) continue;
    ++depth_frames;
    samples = samples && fr.depth_width == 4 && fr.depth_height == 2 && fr.depth_confidence.size() == 8 &&
              fr.depth_confidence[3] == 200 && std::fabs(fr.depth[5] - 1.5f) < 0.006f;
  }
  EXPECT(ds.has_masks && ds.has_depth && depth_frames == 3 && samples,
         "capture: masks %d, depth %d, %zu frames with depth (want 3), samples in metres %d", int(ds.has_masks),
         int(ds.has_depth), depth_frames, int(samples));
  gs::InitOptions init;
  init.max_gaussians = 1000;
  gs::InitStats st;
  const gs::SplatScene scene = gs::initialize_gaussians(ds, init, &st);
  // Frames 2 and 4 train with depth (frame 0 is held out): 8 samples each.
  EXPECT(st.source == "depth" && st.raw_points == 16 && scene.count == 16, "depth init: %s, %zu points, %u gaussians",
         st.source.c_str(), st.raw_points, scene.count);
  // The cameras sit at z = 2 looking down -z, so a sample d metres away lands at z = 2 - d.
  bool placed = true;
  for (uint32_t i = 0; i < scene.count; ++i) {
    const float z = scene.positions[i * 3 + 2];
    placed = placed && (std::fabs(z - (2.0f - 1.502f)) < 1e-3f || std::fabs(z - (2.0f - 1.504f)) < 1e-3f);
  }
  EXPECT(placed, "depth samples must land at z = 2 - depth");
}

void test_merging_and_scales() {
  // A 40 x 40 x 10 lattice (spacing 0.05, under 2% of its extent, so scales
  // are not clamped) given as the sparse point cloud.
  gs::Dataset ds = gs::make_orbit_dataset(4, 16, 16, 3.0f, 0.0f, 60.0f);
  for (int x = 0; x < 40; ++x)
    for (int y = 0; y < 40; ++y)
      for (int z = 0; z < 10; ++z) {
        ds.points.insert(ds.points.end(), {0.05f * x, 0.05f * y, 0.05f * z});
        ds.point_colors.insert(ds.point_colors.end(), {uint8_t(5 * x), uint8_t(5 * y), uint8_t(20 * z)});
      }
  gs::InitOptions o;
  gs::InitStats st;
  o.max_gaussians = 1000;  // over budget: merged per voxel
  const gs::SplatScene merged = gs::initialize_gaussians(ds, o, &st);
  EXPECT(st.source == "points" && st.raw_points == 16000 && st.voxel_size > 0.0f && merged.count > 100 &&
             merged.count <= 1000,
         "merging: %s, %zu raw, voxel %g, %u gaussians", st.source.c_str(), st.raw_points, st.voxel_size,
         merged.count);
  o.max_gaussians = 20000;  // under budget: every point, scale = lattice spacing
  const gs::SplatScene kept = gs::initialize_gaussians(ds, o, &st);
  size_t off = 0;
  for (uint32_t i = 0; i < kept.count; ++i) off += std::fabs(std::exp(kept.log_scales[i * 3]) - 0.05f) > 1e-3f;
  EXPECT(kept.count == 16000 && st.voxel_size == 0.0f && off == 0, "kept %u, %zu scales off the spacing", kept.count,
         off);
  // Morton order: consecutive gaussians are neighbours in space.
  double path = 0.0;
  for (uint32_t i = 1; i < kept.count; ++i) {
    double d2 = 0.0;
    for (int c = 0; c < 3; ++c) d2 += std::pow(kept.positions[i * 3 + c] - kept.positions[(i - 1) * 3 + c], 2);
    path += std::sqrt(d2);
  }
  EXPECT(path / (kept.count - 1) < 0.2, "mean step between consecutive gaussians %.3f", path / (kept.count - 1));
  EXPECT(std::fabs(kept.opacity_logits[0] - std::log(0.2f / 0.8f)) < 1e-6f, "initial opacity 0.2");
  // DC from the colour: colour = 0.2820948 dc + 0.5.
  const float c0 = 0.5f + 0.28209479f * kept.sh_dc[0];
  EXPECT(c0 >= 0.0f && c0 <= 1.0f, "DC gives a colour in [0, 1]: %g", c0);
}

}  // namespace

int main() {
  try {
    test_plain_frames();
    test_capture_with_depth();
    test_merging_and_scales();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_dataset: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
