// gs::SplatRenderer on the GPU, restating the reference's viewer tests
// (engine/tests/test_splats.cpp:188-296): a single splat's falloff against
// the closed form opacity exp(-d^2 / 2 (sigma^2 + 0.3)), depth order, the
// Y-flip equivalence with SH parity signs, and a 20000-splat synthetic
// scene's visible count and coverage.  Also SPZ files decoded on the device
// (v2 and v3, SH degrees 0-4) against the host decoder, value for value and
// image for image.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <random>
#include <stdexcept>
#include <vector>

#include <zlib.h>

#include "gs/camera.h"
#include "gs/device.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"
#include "gs_hip.h"
#include "test_util.h"

namespace {

constexpr float kShC0 = 0.28209479177387814f;

// A scene of round splats (one scale on every axis, no rotation) in Y-up
// axes; each is (x, y, z, size, r, g, b, opacity).
gs::SplatScene round_splats(std::initializer_list<std::array<float, 8>> splats) {
  gs::SplatScene s;
  s.convention = gs::AxisConvention::RUB;
  s.resize(uint32_t(splats.size()), 0);
  uint32_t i = 0;
  for (const auto& sp : splats) {
    std::copy(sp.begin(), sp.begin() + 3, s.positions.begin() + 3 * i);
    std::fill(s.log_scales.begin() + 3 * i, s.log_scales.begin() + 3 * i + 3, std::log(sp[3]));
    for (int c = 0; c < 3; ++c) s.sh_dc[3 * i + c] = (sp[4 + c] - 0.5f) / kShC0;  // colour = 0.5 + C0 dc
    s.rotations[4 * i] = 1.0f;
    s.opacity_logits[i] = -std::log(1.0f / sp[7] - 1.0f);
    ++i;
  }
  return s;
}

// Looking at the origin along -z from `distance` (the orbit's start angles).
gs::OrbitCamera facing_origin(float distance) {
  gs::OrbitCamera cam;
  cam.distance = distance;
  return cam;
}

std::vector<uint8_t> render(gs::Device& dev, const gs::SplatScene& scene, const gs::OrbitCamera& orbit, uint32_t w,
                            uint32_t h, uint32_t* visible = nullptr) {
  auto gpu = gs::upload_scene(dev, scene, w, h);
  gs::SplatRenderer r(dev);
  r.attach(gpu.get());
  const float clear[3] = {0.0f, 0.0f, 0.0f};
  std::vector<uint8_t> img = r.render(orbit.camera(w, h), clear);
  if (visible != nullptr) *visible = r.visible_count();
  return img;
}

const uint8_t* pixel(const std::vector<uint8_t>& img, uint32_t w, uint32_t x, uint32_t y) {
  return &img[(size_t(y) * w + x) * 4];
}

// One round red splat of world size 0.1 at the origin, 3 units away: along a
// diagonal from the centre, red follows opacity exp(-d^2 / 2 (sigma^2 + 0.3))
// with sigma its projected size in pixels (within 6 of 255); green and blue
// stay near zero; far corners keep the black clear colour.
void test_single_splat(gs::Device& dev) {
  const uint32_t n = 64;
  const float o = 0.95f;
  uint32_t visible = 0;
  const auto img = render(dev, round_splats({{0, 0, 0, 0.1f, 1, 0, 0, o}}), facing_origin(3.0f), n, n, &visible);
  EXPECT(visible == 1, "one splat drawn, not %u", visible);
  // The orbit camera's 50 degree field of view spans the short side.
  const float f = 32.0f / std::tan(0.5f * 50.0f * 3.14159265f / 180.0f);
  const float var = (f * 0.1f / 3.0f) * (f * 0.1f / 3.0f) + 0.3f;
  for (int k = 0; k <= 3; ++k) {
    const float d2 = 2.0f * (k + 0.5f) * (k + 0.5f);  // pixel (32 + k, 32 + k), centre offsets k + 0.5
    const float want = 255.0f * std::min(o * std::exp(-0.5f * d2 / var), 0.99f);
    const uint8_t* p = pixel(img, n, 32 + k, 32 + k);
    EXPECT(std::fabs(float(p[0]) - want) <= 6.0f, "diagonal step %d: red %u, closed form %.1f", k, p[0], want);
    EXPECT(int(p[1]) + int(p[2]) <= 4, "diagonal step %d: green %u / blue %u in a red splat", k, p[1], p[2]);
  }
  for (uint32_t c : {1u, 62u}) {
    const uint8_t* p = pixel(img, n, c, c);
    EXPECT(int(p[0]) + int(p[1]) + int(p[2]) == 0, "corner (%u, %u) is not the clear colour", c, c);
  }
  std::printf("single splat: centre red %u\n", pixel(img, n, 32, 32)[0]);
}

// Two opaque splats on the view axis, red and blue: the nearer one (larger z:
// the camera looks down -z from z = 3) owns the centre pixel.
void test_depth_order(gs::Device& dev) {
  for (float red_z : {0.5f, -0.5f}) {
    const auto img = render(dev, round_splats({{0, 0, red_z, 0.1f, 1, 0, 0, 0.9f}, {0, 0, -red_z, 0.1f, 0, 0, 1, 0.9f}}),
                            facing_origin(3.0f), 64, 64);
    const uint8_t* p = pixel(img, 64, 32, 32);
    const int near = red_z > 0.0f ? p[0] : p[2], far = red_z > 0.0f ? p[2] : p[0];
    EXPECT(near > 180 && far < 60, "%s splat nearer: centre r %u b %u", red_z > 0.0f ? "red" : "blue", p[0], p[2]);
  }
}

// Y-down data drawn with the display flip equals the Y-up data drawn without.
// The reference's flip-equivalence criterion (test_splats.cpp:233-262): one
// scene stored Y-up (RUB) and the same scene turned into the Y-down (RDF)
// convention must render alike, the second through the flipped orbit camera.
// The RDF copy comes from gs::convert_convention (the half turn about x on
// positions and rotations, and each SH coefficient's parity sign), over a
// seeded scene with every SH degree-3 coefficient set.
void test_flip_equivalence(gs::Device& dev) {
  gs::SplatScene up = gs::make_synthetic_scene(600, 11);
  up.convention = gs::AxisConvention::RUB;
  {
    gs::SplatScene with_rest;
    with_rest.convention = gs::AxisConvention::RUB;
    with_rest.resize(up.count, 3);
    with_rest.positions = up.positions;
    with_rest.log_scales = up.log_scales;
    with_rest.rotations = up.rotations;
    with_rest.opacity_logits = up.opacity_logits;
    with_rest.sh_dc = up.sh_dc;
    std::mt19937_64 gen(12);
    std::normal_distribution<float> n(0.0f, 0.25f);
    for (float& v : with_rest.sh_rest) v = n(gen);
    up = std::move(with_rest);
  }
  gs::SplatScene down = up;
  gs::convert_convention(down, gs::AxisConvention::RDF);
  gs::OrbitCamera orbit;
  orbit.reset(gs::compute_bounds(up), false);
  orbit.yaw += 0.6f;
  orbit.pitch = 0.35f;
  const std::vector<uint8_t> a = render(dev, up, orbit, 112, 80);
  orbit.flip_y = true;
  const std::vector<uint8_t> b = render(dev, down, orbit, 112, 80);
  size_t covered = 0;
  int worst = 0;
  for (size_t p = 0; p < a.size(); p += 4) {
    covered += std::max({a[p], a[p + 1], a[p + 2]}) > 12;  // drawn over the black clear colour
    for (int c = 0; c < 4; ++c) worst = std::max(worst, std::abs(int(a[p + c]) - int(b[p + c])));
  }
  EXPECT(covered > a.size() / 4 / 10, "flip: only %zu covered pixels", covered);
  EXPECT(worst <= 3, "flip: the two conventions' images differ by %d levels", worst);
  std::printf("flip equivalence: %zu covered pixels, largest difference %d levels\n", covered, worst);
}

// The reference's synthetic-scene criterion (test_splats.cpp:264-291): a
// 20000-splat demo scene framed by the orbit camera's reset is mostly drawn
// and covers a fair part of the image; here also: drawn twice, the image is
// the same byte for byte.
void test_synthetic_render(gs::Device& dev) {
  const gs::SplatScene scene = gs::make_synthetic_scene(20000);
  gs::OrbitCamera orbit;
  orbit.reset(gs::compute_bounds(scene), false);
  uint32_t drawn = 0, drawn_again = 0;
  const std::vector<uint8_t> first = render(dev, scene, orbit, 160, 120, &drawn);
  const std::vector<uint8_t> second = render(dev, scene, orbit, 160, 120, &drawn_again);
  size_t covered = 0;
  for (size_t p = 0; p < first.size(); p += 4) covered += std::max({first[p], first[p + 1], first[p + 2]}) > 12;
  EXPECT(drawn * 4 > 20000 * 3, "synthetic scene: %u of 20000 splats drawn", drawn);
  EXPECT(covered * 10 > first.size() / 4, "synthetic scene covers %zu of %zu pixels", covered, first.size() / 4);
  EXPECT(first == second && drawn == drawn_again, "synthetic scene: two renders differ");
  std::printf("synthetic scene: %u splats drawn, %zu pixels covered\n", drawn, covered);
}

void test_size_check(gs::Device& dev) {
  auto gpu = gs::upload_scene(dev, round_splats({{0, 0, 0, 0.1f, 1, 0, 0, 0.5f}}), 64, 64);
  gs::SplatRenderer r(dev);
  r.attach(gpu.get());
  const float clear[3] = {0, 0, 0};
  bool threw = false;
  try {
    r.render(facing_origin(3.0f).camera(32, 32), clear);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT(threw, "a camera of another size should be rejected");
}

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

// An SPZ file of n splats with random attribute bytes (positions within a
// few units at 12 fractional bits, every byte value of alpha, colour, scale,
// rotation and SH reachable).
std::vector<uint8_t> random_spz(uint32_t n, int degree, uint32_t version, uint32_t seed) {
  const int coeffs = degree <= 0 ? 0 : (degree + 1) * (degree + 1) - 1;
  const size_t rot = version >= 3 ? 4 : 3;
  std::vector<uint8_t> raw(16 + size_t(n) * (9 + 1 + 3 + 3 + rot + size_t(coeffs) * 3));
  const uint32_t magic = 0x5053474eu;
  std::memcpy(&raw[0], &magic, 4);
  std::memcpy(&raw[4], &version, 4);
  std::memcpy(&raw[8], &n, 4);
  raw[12] = uint8_t(degree);
  raw[13] = 12;
  raw[14] = uint8_t(seed & 1u);  // antialiased flag
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> byte(0, 255), pos(-3 * 4096, 3 * 4096);
  size_t o = 16;
  for (uint32_t i = 0; i < n * 3; ++i, o += 3) {
    const int32_t v = pos(rng);
    raw[o] = uint8_t(v), raw[o + 1] = uint8_t(v >> 8), raw[o + 2] = uint8_t(v >> 16);
  }
  for (; o < raw.size(); ++o) raw[o] = uint8_t(byte(rng));
  return gzip(raw);
}

void test_spz_on_device(gs::Device& dev) {
  for (uint32_t version : {2u, 3u}) {
    for (int degree : {0, 1, 3, 4}) {
      const uint32_t n = 3000;
      const std::vector<uint8_t> file = random_spz(n, degree, version, 100 + version * 10 + uint32_t(degree));
      const gs::SplatScene host = gs::load_spz(file.data(), file.size());
      gs::GpuScene gpu(dev, file.data(), file.size(), 128, 96);
      EXPECT(gpu.count() == n && gpu.sh_degree() == std::min(degree, 3) && gpu.antialiased() == host.antialiased,
             "spz header on the device");
      const int coeffs = gs::sh_rest_coeffs(gpu.sh_degree());
      std::vector<float> p(size_t(n) * 16), rest(size_t(n) * std::max(coeffs, 1) * 3);
      gs::check_status(gs_hip_workspace_store(gpu.workspace(), p.data(), coeffs > 0 ? rest.data() : nullptr, n));
      size_t bad = 0;
      for (uint32_t i = 0; i < n; ++i) {
        const float* d = &p[size_t(i) * 16];
        for (int k = 0; k < 3; ++k) {
          bad += std::memcmp(&d[k], &host.positions[i * 3 + k], 4) != 0;
          bad += std::memcmp(&d[4 + k], &host.log_scales[i * 3 + k], 4) != 0;
          bad += std::memcmp(&d[12 + k], &host.sh_dc[i * 3 + k], 4) != 0;
        }
        bad += std::memcmp(&d[3], &host.opacity_logits[i], 4) != 0;
        bad += d[7] != 1.0f;
        for (int k = 0; k < 4; ++k) bad += std::memcmp(&d[8 + k], &host.rotations[i * 4 + k], 4) != 0;
      }
      for (size_t k = 0; k < host.sh_rest.size(); ++k) bad += std::memcmp(&rest[k], &host.sh_rest[k], 4) != 0;
      EXPECT(bad == 0, "spz v%u degree %d: %zu values differ from the host decoder", version, degree, bad);
      // and the same image as the host-decoded scene uploaded
      gs::OrbitCamera orbit;
      orbit.reset(gs::compute_bounds(host), false);
      auto uploaded = gs::upload_scene(dev, host, 128, 96);
      gs::SplatRenderer a(dev), b(dev);
      a.attach(uploaded.get());
      b.attach(&gpu);
      const float clear[3] = {0.1f, 0.1f, 0.1f};
      const auto ia = a.render(orbit.camera(128, 96), clear), ib = b.render(orbit.camera(128, 96), clear);
      EXPECT(ia == ib && a.visible_count() == b.visible_count() && a.visible_count() > 100,
             "spz v%u degree %d: device-decoded image differs (%u / %u visible)", version, degree,
             a.visible_count(), b.visible_count());
    }
  }
  std::printf("spz on the device: v2/v3, degrees 0-4, equal to the host decoder\n");
}

}  // namespace

int main() {
  try {
    gs::Device dev(0);
    std::printf("device: %s\n", dev.name().c_str());
    std::fflush(stdout);
    test_single_splat(dev);
    test_depth_order(dev);
    test_flip_equivalence(dev);
    test_synthetic_render(dev);
    test_size_check(dev);
    test_spz_on_device(dev);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_renderer: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
