// gs::SplatRenderer on the GPU, restating the reference's viewer tests
// (engine/tests/test_splats.cpp:188-296): a single splat's falloff against
// the closed form opacity exp(-d^2 / 2 (sigma^2 + 0.3)), depth order, the
// Y-flip equivalence with SH parity signs, and a 20000-splat synthetic
// scene's visible count and coverage.  Also SPZ files decoded on the device
// (v2 and v3, SH degrees 0-4) against the host decoder, value for value and
// image for image.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

struct TestSplat {
  float pos[3];
  float scale;
  float rgb[3];
  float opacity;
};

gs::SplatScene splats_scene(const std::vector<TestSplat>& splats) {
  gs::SplatScene s;
  s.convention = gs::AxisConvention::RUB;
  s.resize(uint32_t(splats.size()), 0);
  for (size_t i = 0; i < splats.size(); ++i) {
    const TestSplat& t = splats[i];
    for (int k = 0; k < 3; ++k) {
      s.positions[i * 3 + k] = t.pos[k];
      s.log_scales[i * 3 + k] = std::log(t.scale);
      s.sh_dc[i * 3 + k] = (t.rgb[k] - 0.5f) / kShC0;
    }
    s.rotations[i * 4] = 1.0f;
    s.opacity_logits[i] = std::log(t.opacity / (1.0f - t.opacity));
  }
  return s;
}

gs::OrbitCamera front_camera() {
  gs::OrbitCamera orbit;
  orbit.yaw = 0.0f;
  orbit.pitch = 0.0f;
  orbit.distance = 3.0f;
  return orbit;
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

void test_single_splat(gs::Device& dev) {
  const uint32_t w = 64, h = 64;
  const float opacity = 0.95f;
  uint32_t visible = 0;
  const auto img = render(dev, splats_scene({{{0, 0, 0}, 0.1f, {1, 0, 0}, opacity}}), front_camera(), w, h, &visible);
  EXPECT(visible == 1, "single splat: visible %u", visible);
  const float focal = 0.5f * 64.0f / std::tan(25.0f * 3.14159265f / 180.0f);
  const float sigma = focal * 0.1f / 3.0f;
  const float var = sigma * sigma + 0.3f;
  for (int dx : {0, 2, 4}) {
    const float px = dx + 0.5f, py = 0.5f;
    const float alpha = std::min(opacity * std::exp(-0.5f * (px * px + py * py) / var), 0.99f);
    const uint8_t* p = pixel(img, w, 32 + dx, 32);
    EXPECT(std::fabs(p[0] - 255.0f * alpha) <= 6.0f, "falloff at dx=%d: red %u, expected %.1f", dx, p[0],
           255.0f * alpha);
    EXPECT(p[1] <= 2 && p[2] <= 2, "single splat should be pure red (g %u b %u)", p[1], p[2]);
  }
  const uint8_t* corner = pixel(img, w, 2, 2);
  EXPECT(corner[0] == 0 && corner[1] == 0 && corner[2] == 0, "background should stay black");
  std::printf("single splat: centre red %u\n", pixel(img, w, 32, 32)[0]);
}

void test_depth_order(gs::Device& dev) {
  for (bool red_in_front : {true, false}) {
    const float red_z = red_in_front ? 0.5f : -0.5f;  // the camera sits at +z: larger z is nearer
    const auto img = render(dev,
                            splats_scene({{{0, 0, red_z}, 0.1f, {1, 0, 0}, 0.9f},
                                          {{0, 0, -red_z}, 0.1f, {0, 0, 1}, 0.9f}}),
                            front_camera(), 64, 64);
    const uint8_t* p = pixel(img, 64, 32, 32);
    if (red_in_front) {
      EXPECT(p[0] > 180 && p[2] < 60, "red in front: got r %u b %u", p[0], p[2]);
    } else {
      EXPECT(p[2] > 180 && p[0] < 60, "blue in front: got r %u b %u", p[0], p[2]);
    }
  }
}

// Y-down data drawn with the display flip equals the Y-up data drawn without.
void test_flip_equivalence(gs::Device& dev) {
  std::mt19937 rng(3);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  gs::SplatScene up;
  up.convention = gs::AxisConvention::RUB;
  up.resize(400, 3);
  for (uint32_t i = 0; i < up.count; ++i) {
    for (int k = 0; k < 3; ++k) {
      up.positions[i * 3 + k] = 0.8f * u(rng);
      up.log_scales[i * 3 + k] = std::log(0.03f + 0.05f * std::fabs(u(rng)));
      up.sh_dc[i * 3 + k] = 1.5f * u(rng);
    }
    for (int k = 0; k < 4; ++k) up.rotations[i * 4 + k] = u(rng);
    up.opacity_logits[i] = 2.0f * u(rng);
    for (int j = 0; j < 45; ++j) up.sh_rest[i * 45 + j] = 0.3f * u(rng);
  }
  // Half turn about X: (x, y, z) -> (x, -y, -z), quaternion (w, x, y, z) -> (w, x, -y, -z); each SH basis
  // function takes the sign of its parity in y and z.
  const float sh_sign[15] = {-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
  gs::SplatScene down = up;
  down.convention = gs::AxisConvention::RDF;
  for (uint32_t i = 0; i < down.count; ++i) {
    down.positions[i * 3 + 1] *= -1.0f;
    down.positions[i * 3 + 2] *= -1.0f;
    down.rotations[i * 4 + 2] *= -1.0f;
    down.rotations[i * 4 + 3] *= -1.0f;
    for (int j = 0; j < 15; ++j) {
      for (int c = 0; c < 3; ++c) down.sh_rest[(size_t(i) * 15 + j) * 3 + c] *= sh_sign[j];
    }
  }
  gs::OrbitCamera orbit;
  orbit.yaw = 0.7f;
  orbit.pitch = 0.4f;
  orbit.distance = 3.0f;
  const auto img_up = render(dev, up, orbit, 96, 96);
  orbit.flip_y = true;
  const auto img_down = render(dev, down, orbit, 96, 96);
  size_t lit = 0;
  for (size_t i = 0; i < img_up.size(); i += 4) lit += (img_up[i] + img_up[i + 1] + img_up[i + 2]) > 30;
  EXPECT(lit > 500, "flip equivalence: scene barely visible (%zu lit pixels)", lit);
  int worst = 0;
  for (size_t i = 0; i < img_up.size(); ++i) worst = std::max(worst, std::abs(int(img_up[i]) - int(img_down[i])));
  EXPECT(worst <= 3, "flip equivalence: max channel difference %d", worst);
  std::printf("flip equivalence: %zu lit pixels, max difference %d\n", lit, worst);
}

void test_synthetic_render(gs::Device& dev) {
  const gs::SplatScene scene = gs::make_synthetic_scene(20000);
  gs::OrbitCamera orbit;
  orbit.reset(gs::compute_bounds(scene), false);
  uint32_t visible = 0;
  const auto img = render(dev, scene, orbit, 160, 120, &visible);
  EXPECT(visible > 15000, "synthetic: only %u of 20000 visible", visible);
  size_t lit = 0;
  for (size_t i = 0; i < img.size(); i += 4) lit += (img[i] + img[i + 1] + img[i + 2]) > 30;
  EXPECT(lit > img.size() / 4 / 10, "synthetic: only %zu lit pixels", lit);
  std::printf("synthetic: %u visible, %zu lit pixels\n", visible, lit);
}

void test_size_check(gs::Device& dev) {
  auto gpu = gs::upload_scene(dev, splats_scene({{{0, 0, 0}, 0.1f, {1, 0, 0}, 0.5f}}), 64, 64);
  gs::SplatRenderer r(dev);
  r.attach(gpu.get());
  const float clear[3] = {0, 0, 0};
  bool threw = false;
  try {
    r.render(front_camera().camera(32, 32), clear);
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
