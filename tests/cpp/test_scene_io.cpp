// Scene I/O and camera conventions of the C++ API, without a GPU:
// PLY round trip (bit-equal), ASCII PLY rejected, the reference's SPZ v2
// known answers (engine/tests/test_splats.cpp:34-130), SPZ v3 quaternions,
// v4 rejected by name, the axis flip and dataset cameras; the viewer's fp16
// rounding of SH rest against the compiler's _Float16 conversion, over every
// fp16 value and the midpoints between neighbours.
#include <zlib.h>

#include <cmath>
#include <cstring>
#include <initializer_list>
#include <random>
#include <stdexcept>
#include <string>

#include "gs/camera.h"
#include "gs/dataset.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"
#include "test_util.h"

namespace {

std::vector<uint8_t> gzip(const std::vector<uint8_t>& data) {
  z_stream zs{};
  deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 16 + MAX_WBITS, 8, Z_DEFAULT_STRATEGY);
  std::vector<uint8_t> out(deflateBound(&zs, data.size()) + 64);
  zs.next_in = const_cast<Bytef*>(data.data());
  zs.avail_in = uInt(data.size());
  zs.next_out = out.data();
  zs.avail_out = uInt(out.size());
  deflate(&zs, Z_FINISH);
  out.resize(zs.total_out);
  deflateEnd(&zs);
  return out;
}

gs::SplatScene random_scene(uint32_t n, int degree, uint32_t seed) {
  gs::SplatScene s;
  s.resize(n, degree);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  for (auto* v : {&s.positions, &s.log_scales, &s.rotations, &s.opacity_logits, &s.sh_dc, &s.sh_rest})
    for (float& x : *v) x = u(rng);
  return s;
}

void test_ply_roundtrip() {
  for (int degree = 0; degree <= 3; ++degree) {
    const gs::SplatScene s = random_scene(257, degree, 3 + degree);
    const std::vector<uint8_t> bytes = gs::save_ply(s);
    const gs::SplatScene r = gs::load_ply(bytes.data(), bytes.size());
    EXPECT(r.count == s.count && r.sh_degree == degree, "ply header: %u %d", r.count, r.sh_degree);
    EXPECT(r.positions == s.positions && r.log_scales == s.log_scales && r.rotations == s.rotations &&
               r.opacity_logits == s.opacity_logits && r.sh_dc == s.sh_dc && r.sh_rest == s.sh_rest,
           "ply round trip differs at degree %d", degree);
    const gs::SplatScene d = gs::load_splat_file(bytes.data(), bytes.size());
    EXPECT(d.count == s.count, "load_splat_file detects PLY");
  }
  bool threw = false;
  const std::string ascii = "ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n0\n";
  try {
    gs::load_ply(reinterpret_cast<const uint8_t*>(ascii.data()), ascii.size());
  } catch (const std::runtime_error&) {
    threw = true;
  }
  EXPECT(threw, "ascii PLY must be rejected");
}

// An SPZ payload before compression: the 16-byte header (magic "NGSP",
// version, count, SH degree, 12 fractional position bits, flags, reserved),
// then whatever the test appends, little-endian.
struct SpzPayload {
  std::vector<uint8_t> bytes;
  SpzPayload(uint32_t version, uint32_t count, int degree) {
    put(0x5053474e, 4);
    put(version, 4);
    put(count, 4);
    put(uint32_t(degree) | 12u << 8, 4);
  }
  void put(uint32_t v, int width) {
    for (int i = 0; i < width; ++i) bytes.push_back(uint8_t(v >> (8 * i)));
  }
  void put_each(std::initializer_list<int> values) {
    for (int v : values) bytes.push_back(uint8_t(v));
  }
};

void test_spz_v2() {
  SpzPayload spz(2, 2, 1);
  const float pos[2][3] = {{1.5f, -2.25f, 0.125f}, {-0.5f, 0.0f, 3.0f}};
  for (const auto& p : pos) {
    for (float v : p) spz.put(uint32_t(int32_t(std::lround(v * 4096.0f))), 3);  // 24-bit fixed point
  }
  spz.put_each({255, 128});                             // alphas
  spz.put_each({64, 84, 104, 124, 144, 164});           // colours
  spz.put_each({100, 101, 102, 103, 104, 105});         // log scales
  spz.put_each({128, 128, 128, 255, 128, 128});         // quaternions (x y z): identity, then a half turn about x
  for (int i = 0; i < 18; ++i) spz.put(128 + i, 1);     // SH rest, 9 per gaussian
  const std::vector<uint8_t> file = gzip(spz.bytes);
  const gs::SplatScene s = gs::load_spz(file.data(), file.size());
  EXPECT(s.count == 2 && s.sh_degree == 1, "spz header %u %d", s.count, s.sh_degree);
  EXPECT(s.convention == gs::AxisConvention::RUB, "spz is RUB");
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 3; ++k) EXPECT(std::fabs(s.positions[i * 3 + k] - pos[i][k]) < 1e-3f, "spz position");
  EXPECT(s.opacity_logits[0] > 5.0f, "alpha 255 nearly opaque");
  EXPECT(std::fabs(s.opacity_logits[1] - std::log(128.0f / 127.0f)) < 1e-3f, "alpha 128");
  EXPECT(std::fabs(s.sh_dc[0] - (64 / 255.0f - 0.5f) / 0.15f) < 1e-5f, "spz colour");
  EXPECT(std::fabs(s.log_scales[0] - (100 / 16.0f - 10.0f)) < 1e-5f, "spz scale");
  EXPECT(std::fabs(s.rotations[0] - 1.0f) < 0.01f, "identity rotation w %f", s.rotations[0]);
  EXPECT(std::fabs(s.rotations[5] - 1.0f) < 0.01f, "x rotation x %f", s.rotations[5]);
  EXPECT(std::fabs(s.sh_rest[5] - 5.0f / 128.0f) < 1e-6f, "spz SH layout");
  // A version 4 header, uncompressed, is refused by name.
  const SpzPayload v4(4, 0, 0);
  std::string message;
  try {
    gs::load_splat_file(v4.bytes.data(), 8);
  } catch (const std::runtime_error& e) {
    message = e.what();
  }
  EXPECT(message.find("v4") != std::string::npos, "spz v4 rejected by name (\"%s\")", message.c_str());
}

// Smallest-three packing of (w, x, y, z) = (0.8, 0.6, 0, 0): the largest
// component, w, is index 3 of (x, y, z, w) in the top two bits; the other
// three follow as 9-bit magnitudes over 1/sqrt(2) with a sign bit, 10 bits
// each, the last (z) in the low bits.
void test_spz_v3_quaternion() {
  SpzPayload spz(3, 1, 0);
  for (int i = 0; i < 9; ++i) spz.put(0, 1);  // position at the origin
  spz.put_each({200, 128, 128, 128, 100, 100, 100});  // alpha, colour, log scales
  const auto magnitude = [](float v) { return uint32_t(std::lround(std::fabs(v) / 0.70710678f * 511.0f)); };
  const uint32_t packed = 3u << 30 | magnitude(0.6f) << 20 | magnitude(0.0f) << 10 | magnitude(0.0f);
  spz.put(packed, 4);
  const std::vector<uint8_t> file = gzip(spz.bytes);
  const gs::SplatScene s = gs::load_spz(file.data(), file.size());
  EXPECT(std::fabs(s.rotations[0] - 0.8f) < 2e-3f && std::fabs(s.rotations[1] - 0.6f) < 2e-3f,
         "v3 quaternion (%f %f %f %f)", s.rotations[0], s.rotations[1], s.rotations[2], s.rotations[3]);
}

void test_convention_flip() {
  gs::SplatScene s = random_scene(50, 3, 9);
  const gs::SplatScene orig = s;
  gs::convert_convention(s, gs::AxisConvention::RUB);
  EXPECT(s.convention == gs::AxisConvention::RUB, "flip sets convention");
  EXPECT(s.positions[1] == -orig.positions[1] && s.positions[2] == -orig.positions[2], "flip negates y, z");
  gs::convert_convention(s, gs::AxisConvention::RDF);
  // Two half turns: positions and SH back; the quaternion is negated (same rotation).
  EXPECT(s.positions == orig.positions && s.sh_rest == orig.sh_rest, "double flip restores");
  for (size_t i = 0; i < s.rotations.size(); ++i) EXPECT(s.rotations[i] == -orig.rotations[i], "quat sign");
}

void test_dataset_camera() {
  gs::Dataset ds = gs::make_orbit_dataset(8, 64, 48, 3.0f, 0.5f, 60.0f, 4);
  EXPECT(ds.holdout_count() == 2 && ds.train_count() == 6, "holdout every 4th");
  EXPECT(std::fabs(ds.camera_extent - 1.1f * 3.0f) < 1e-4f, "extent %f", ds.camera_extent);
  for (const gs::DatasetFrame& f : ds.frames) {
    const gs::Camera c = gs::dataset_camera(ds, f);
    // The origin projects to the image centre, in front of the camera.
    const float z = c.translation[2];
    EXPECT(z > 2.9f && z < 3.1f, "depth of the origin %f", z);
    EXPECT(std::fabs(c.translation[0]) < 1e-4f && std::fabs(c.translation[1]) < 1e-4f, "origin on axis");
    // Rows orthonormal, y down: scene +Y maps to negative camera y.
    EXPECT(c.rotation[4] < 0.0f, "y down");
  }
}

void test_round_to_half() {
  int bad = 0;
  float prev = -1.0f;
  for (uint32_t b = 0; b < 0x7c00u; ++b) {  // every finite non-negative fp16
    _Float16 h;
    const uint16_t bits = uint16_t(b);
    std::memcpy(&h, &bits, 2);
    const float v = float(h);
    for (float x : {v, -v}) bad += gs::round_to_half(x) != float(_Float16(x));
    if (prev >= 0.0f) {
      const float mid = float((double(prev) + double(v)) / 2.0);  // exact in fp32
      for (float x : {mid, std::nextafter(mid, 0.0f), std::nextafter(mid, 1e30f), -mid}) {
        bad += gs::round_to_half(x) != float(_Float16(x));
      }
    }
    prev = v;
  }
  for (float x : {65519.0f, 65520.0f, 1e6f, 1e-9f, -65520.0f}) bad += gs::round_to_half(x) != float(_Float16(x));
  EXPECT(bad == 0, "round_to_half: %d values differ from _Float16", bad);
}

}  // namespace

int main() {
  test_ply_roundtrip();
  test_spz_v2();
  test_spz_v3_quaternion();
  test_convention_flip();
  test_dataset_camera();
  test_round_to_half();
  if (test::failures() == 0) std::printf("test_scene_io: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
