// gs::Trainer on the GPU, restating the reference's trainer tests
// (engine/tests/test_trainer.cpp): pair order (197-233), frozen-rate
// gradients and parameters, render size check, densification at capacity
// (866-888) and fitting a known scene (890-897; train_selftest.cpp:692-739),
// one view per optimizer step and four (batched, autotuned).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <tuple>

#include "gs/dataset.h"
#include "gs/device.h"
#include "gs/splat_scene.h"
#include "gs/trainer.h"
#include "test_util.h"

namespace {

// Orbit at radius 3 with a wobbling height, fx = fy = 0.9 W, random-noise
// targets (make_test_dataset, train_selftest.cpp:285-312).
gs::Dataset test_dataset(uint32_t w, uint32_t h, size_t frames, bool masks, uint32_t seed) {
  gs::Dataset ds;
  ds.name = "test";
  ds.width = w;
  ds.height = h;
  ds.fx = ds.fy = 0.9f * w;
  ds.cx = w * 0.5f;
  ds.cy = h * 0.5f;
  ds.has_masks = masks;
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> byte(0, 255);
  for (size_t i = 0; i < frames; ++i) {
    const float a = 6.2831853f * float(i) / float(std::max<size_t>(frames, 1));
    const float eye[3] = {3.0f * std::sin(a), 0.6f + 0.3f * std::sin(3.0f * a), 3.0f * std::cos(a)};
    const float origin[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    const gs::Camera c = gs::look_at(eye, origin, up, w, h, 60.0f);
    gs::DatasetFrame f;
    for (int r = 0; r < 3; ++r) {
      f.c2w[r * 4 + 0] = c.rotation[r];
      f.c2w[r * 4 + 1] = -c.rotation[3 + r];
      f.c2w[r * 4 + 2] = -c.rotation[6 + r];
      f.c2w[r * 4 + 3] = eye[r];
    }
    f.c2w[12] = f.c2w[13] = f.c2w[14] = 0.0f;
    f.c2w[15] = 1.0f;
    f.rgba.resize(size_t(w) * h * 4);
    for (uint8_t& v : f.rgba) v = uint8_t(byte(rng));
    if (!masks)
      for (size_t p = 0; p < size_t(w) * h; ++p) f.rgba[p * 4 + 3] = 255;
    f.holdout = frames > 1 && i % 8 == 0;
    ds.frames.push_back(std::move(f));
  }
  ds.camera_extent = 3.3f;
  return ds;
}

// make_test_scene (train_selftest.cpp:314-332).
gs::SplatScene test_scene(uint32_t n, int degree, uint32_t seed) {
  gs::SplatScene s;
  s.convention = gs::AxisConvention::RUB;
  s.resize(n, degree);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      s.positions[i * 3 + k] = 0.7f * u(rng);
      s.log_scales[i * 3 + k] = std::log(0.08f + 0.06f * u(rng));
      s.sh_dc[i * 3 + k] = 0.8f * u(rng);
    }
    for (int k = 0; k < 4; ++k) s.rotations[i * 4 + k] = u(rng);
    s.rotations[i * 4] += 1.5f;
    s.opacity_logits[i] = 0.6f * u(rng) - 0.2f;
  }
  for (float& v : s.sh_rest) v = 0.25f * u(rng);
  return s;
}

gs::TrainOptions frozen(int degree) {
  gs::TrainOptions o;
  o.iterations = 1;
  o.max_gaussians = 0;
  o.sh_degree = degree;
  o.sh_degree_interval = 0;
  o.masked = false;
  o.random_background = false;
  o.lr_position = o.lr_position_final = o.lr_scale = o.lr_rotation = o.lr_opacity = o.lr_sh_dc = o.lr_sh_rest = 0.0f;
  o.densify_from = o.densify_until = 1000000;
  return o;
}

void test_pair_order(gs::Device& dev) {
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 31), test_scene(400, 1, 37), frozen(1));
  const float bg[3] = {0, 0, 0};
  t.step_on(0, bg);
  const gs::Trainer::DebugSnapshot s = t.debug_snapshot();
  std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> got, want;
  for (size_t k = 0; k < s.pair_tile.size(); ++k) {
    got.emplace_back(s.pair_tile[k], s.depth_key.at(s.pair_gaussian[k]), s.pair_gaussian[k]);
  }
  want = got;
  std::sort(want.begin(), want.end());
  size_t ties = 0;
  for (size_t k = 1; k < want.size(); ++k) {
    ties += std::get<0>(want[k]) == std::get<0>(want[k - 1]) && std::get<1>(want[k]) == std::get<1>(want[k - 1]);
  }
  EXPECT(got.size() > 500, "too few pairs: %zu", got.size());
  EXPECT(got == want, "pairs not in (tile, depth, gaussian) order");
  uint64_t sum = 0;
  for (uint32_t c : s.tiles) sum += c;
  EXPECT(sum == got.size(), "tile counts %llu vs pairs %zu", (unsigned long long)sum, got.size());
  std::printf("  pair order: %zu pairs, %zu ties in (tile, depth)\n", got.size(), ties);
}

void test_frozen_step(gs::Device& dev) {
  const gs::SplatScene scene = test_scene(80, 2, 5);
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 7), scene, frozen(2));
  const float bg[3] = {0.2f, 0.1f, 0.0f};
  t.step_on(0, bg);
  const gs::Trainer::Gradients g = t.first_step_gradients();
  double norm = 0.0;
  for (float v : g.params) norm += std::fabs(v);
  EXPECT(norm > 0.0 && std::isfinite(norm), "gradient norm %g", norm);
  EXPECT(g.sh.size() == size_t(80) * 24, "sh gradient size %zu", g.sh.size());
  const gs::SplatScene after = t.export_scene();
  EXPECT(after.count == 80, "export count %u", after.count);
  bool same = true;
  for (size_t i = 0; i < after.positions.size(); ++i) same = same && after.positions[i] == scene.positions[i];
  EXPECT(same, "frozen rates must leave positions unchanged");
  EXPECT(t.stats().pairs > 0 && t.stats().loss > 0.0, "stats: pairs %llu loss %f",
         (unsigned long long)t.stats().pairs, t.stats().loss);
  bool threw = false;
  try {
    gs::Camera c = t.camera_for(0);
    c.width = 32;
    t.render(c, bg);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT(threw, "render with the wrong size must throw std::invalid_argument");
}

void test_densify_at_capacity(gs::Device& dev) {
  gs::TrainOptions o = frozen(1);
  o.iterations = 3;
  o.max_gaussians = 60;
  o.densify_from = 0;
  o.densify_until = 100;
  o.densify_interval = 1;
  o.densify_grad_threshold = 0.0f;
  o.lr_position = 1e-3f;
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 17), test_scene(40, 1, 19), o);
  const float bg[3] = {0, 0, 0};
  t.step_on(0, bg);
  t.step_on(0, bg);
  EXPECT(t.stats().slots == 60 && t.stats().gaussians == 60, "slots %u alive %u", t.stats().slots,
         t.stats().gaussians);
  EXPECT(std::isfinite(t.stats().loss), "loss finite");
  EXPECT(t.stats().densify_events >= 1, "densified");
}

// views_per_step > 1: the batched step (mean gradient of the views, one launch
// per stage), run for as many optimizer steps as the plain run, with the
// launch shape autotuned on the first views.
void test_training_converges(gs::Device& dev, uint32_t views_per_step) {
  const int degree = 1;
  gs::Dataset ds = test_dataset(96, 72, 24, true, 9);
  const gs::SplatScene truth = test_scene(60, degree, 21);
  {
    gs::Trainer renderer(dev, ds, truth, frozen(degree));
    const float black[3] = {0, 0, 0};
    for (gs::DatasetFrame& f : ds.frames) {
      std::vector<uint8_t> img = renderer.render(gs::dataset_camera(ds, f), black);
      for (size_t p = 0; p < img.size() / 4; ++p) {
        const float a = img[p * 4 + 3] / 255.0f;
        for (int c = 0; c < 3; ++c) f.rgba[p * 4 + c] = a > 0 ? uint8_t(std::min(255.0f, img[p * 4 + c] / a + 0.5f)) : 0;
        f.rgba[p * 4 + 3] = img[p * 4 + 3];
      }
    }
  }
  gs::SplatScene init = truth;
  std::mt19937 rng(4);
  std::normal_distribution<float> noise(0.0f, 1.0f);
  for (float& v : init.positions) v += 0.05f * noise(rng);
  for (float& v : init.sh_dc) v = 0.0f;
  for (float& v : init.sh_rest) v = 0.0f;
  for (float& v : init.opacity_logits) v = -1.0f;
  gs::TrainOptions o;
  o.iterations = 600;
  o.sh_degree = degree;
  o.sh_degree_interval = 200;
  o.densify_from = o.densify_until = 1000000;
  o.max_gaussians = 2000;
  o.views_per_step = views_per_step;
  o.autotune = views_per_step > 1;
  o.iterations *= views_per_step;
  o.sh_degree_interval *= views_per_step;
  gs::Trainer t(dev, ds, init, o);
  const double before = t.evaluate().psnr;
  while (!t.finished()) t.step();
  const gs::EvalResult after = t.evaluate();
  std::printf("  fit known scene, %u view(s) per step: held-out PSNR %.2f -> %.2f dB (%u views), %u optimizer steps, "
              "%.2f ms/step\n", views_per_step, before, after.psnr, after.views, t.stats().optimizer_steps,
              t.stats().step_ms);
  EXPECT(after.psnr > 35.0 && after.psnr > before + 8.0, "PSNR %.2f -> %.2f", before, after.psnr);
  EXPECT(t.stats().step == o.iterations && t.stats().optimizer_steps == 600, "steps %u, optimizer steps %u",
         t.stats().step, t.stats().optimizer_steps);
}

}  // namespace

int main() {
  try {
    gs::Device dev(0);
    std::printf("device: %s\n", dev.name().c_str());
    std::fflush(stdout);
    test_pair_order(dev);
    test_frozen_step(dev);
    test_densify_at_capacity(dev);
    test_training_converges(dev, 1);
    test_training_converges(dev, 4);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_trainer: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
