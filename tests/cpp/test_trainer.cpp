// gs::Trainer on the GPU, restating the reference's trainer tests
// (engine/tests/test_trainer.cpp): pair order (197-233), frozen-rate
// gradients and parameters, render size check, the pose gradient against the
// per-gaussian gradients (404-464), pose recovery (509-536), exposure
// recovery (700-742), held-out alignment, densification at capacity
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

// The PLY rows written on the device (Trainer::export_ply) equal the host
// path, save_ply(export_scene() in RDF), byte for byte, with pruned (free)
// slots among the alive ones (about half are pruned, none spawn) and SH
// degree 2.
void test_export_ply(gs::Device& dev) {
  gs::TrainOptions o = frozen(2);
  o.iterations = 4;
  o.max_gaussians = 400;
  o.densify_from = 0;
  o.densify_until = 100;
  o.densify_interval = 1;
  o.densify_grad_threshold = 1e9f;
  o.prune_opacity = 0.45f;
  o.lr_position = 1e-3f;
  o.lr_sh_rest = 1e-3f;
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 23), test_scene(150, 2, 29), o);
  const float bg[3] = {0, 0, 0};
  t.step_on(0, bg);
  t.step_on(0, bg);
  gs::SplatScene host = t.export_scene();
  gs::convert_convention(host, gs::AxisConvention::RDF);
  const std::vector<uint8_t> want = gs::save_ply(host), got = t.export_ply();
  EXPECT(t.stats().pruned > 0 && host.count < t.stats().slots, "free slots: %u alive of %u slots", host.count,
         t.stats().slots);
  EXPECT(got == want, "export_ply: %zu bytes vs save_ply %zu bytes", got.size(), want.size());
  std::printf("  export_ply: %u gaussians of %u slots, %zu bytes, equal to save_ply\n", host.count, t.stats().slots,
              got.size());
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

// Targets rendered from `scene` at the dataset's own cameras, alpha 255.
void render_targets(gs::Device& dev, gs::Dataset& ds, const gs::SplatScene& scene) {
  gs::Trainer renderer(dev, ds, scene, frozen(scene.sh_degree));
  const float black[3] = {0, 0, 0};
  for (gs::DatasetFrame& f : ds.frames) {
    f.rgba = renderer.render(gs::dataset_camera(ds, f), black);
    for (size_t p = 0; p < f.rgba.size() / 4; ++p) f.rgba[p * 4 + 3] = 255;
  }
}

// Moves frame f's camera as the pose correction does: world -> camera [W | t]
// becomes [E W | E t + d], E the rotation by w; stored back as camera-to-world
// in OpenGL axes.
void nudge_camera(gs::Dataset& ds, size_t f, const double d[3], const double w[3]) {
  const gs::Camera cam = gs::dataset_camera(ds, ds.frames[f]);
  const double theta = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double k[3] = {w[0] / theta, w[1] / theta, w[2] / theta};
  const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
  double E[9], W[9], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      double k2 = 0.0;
      for (int i = 0; i < 3; ++i) k2 += K[r * 3 + i] * K[i * 3 + c];
      E[r * 3 + c] = (r == c ? 1.0 : 0.0) + std::sin(theta) * K[r * 3 + c] + (1.0 - std::cos(theta)) * k2;
    }
  }
  for (int r = 0; r < 3; ++r) {
    t[r] = d[r];
    for (int c = 0; c < 3; ++c) {
      W[r * 3 + c] = 0.0;
      for (int i = 0; i < 3; ++i) W[r * 3 + c] += E[r * 3 + i] * cam.rotation[i * 3 + c];
      t[r] += E[r * 3 + c] * cam.translation[c];
    }
  }
  float* m = ds.frames[f].c2w;
  for (int r = 0; r < 3; ++r) {
    m[r * 4 + 0] = float(W[0 * 3 + r]);
    m[r * 4 + 1] = float(-W[1 * 3 + r]);
    m[r * 4 + 2] = float(-W[2 * 3 + r]);
    m[r * 4 + 3] = float(-(W[0 * 3 + r] * t[0] + W[1 * 3 + r] * t[1] + W[2 * 3 + r] * t[2]));
  }
}

// Distance between camera centres and the angle between their rotations (degrees).
void camera_error(const gs::Camera& a, const gs::Camera& b, double* shift, double* turn) {
  double d2 = 0.0, trace = 0.0;
  for (int k = 0; k < 3; ++k) d2 += double(a.position[k] - b.position[k]) * (a.position[k] - b.position[k]);
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) trace += double(a.rotation[r * 3 + k]) * b.rotation[r * 3 + k];
  *shift = std::sqrt(d2);
  *turn = std::acos(std::clamp((trace - 1.0) / 2.0, -1.0, 1.0)) * 180.0 / M_PI;
}

// Moving the camera is moving every gaussian the other way: dL/dd = W sum
// dL/dp and dL/dw = W sum((p - c) x dL/dp + tau), tau the orientation's share
// from the quaternion gradient (q -> (1, phi / 2) q).  SH degree 0: colour
// does not depend on the view direction, which the pose gradient leaves out.
void test_pose_gradient(gs::Device& dev, float lambda) {
  gs::Dataset ds = test_dataset(64, 48, 2, true, 3);
  gs::SplatScene scene = test_scene(120, 0, 9);
  const float bg[3] = {0.2f, 0.3f, 0.4f};
  gs::TrainOptions o = frozen(0);
  o.masked = true;
  o.lambda_dssim = o.lambda_dssim_end = lambda;
  o.refine_poses = true;
  o.pose_from = 0;
  o.lr_pose_shift = o.lr_pose_turn = 0.0f;  // measure only
  gs::Trainer t(dev, ds, scene, o);
  t.step_on(0, bg);
  const double* g = t.last_pose_gradient();
  const gs::Trainer::Gradients grads = t.first_step_gradients();
  const gs::Camera cam = gs::dataset_camera(ds, ds.frames[0]);
  double sum_d[3] = {0, 0, 0}, sum_w[3] = {0, 0, 0};
  for (uint32_t i = 0; i < scene.count; ++i) {
    const float* gp = &grads.params[size_t(i) * 16];
    double rel[3];
    for (int k = 0; k < 3; ++k) {
      sum_d[k] += gp[k];
      rel[k] = double(scene.positions[i * 3 + k]) - cam.position[k];
    }
    sum_w[0] += rel[1] * gp[2] - rel[2] * gp[1];
    sum_w[1] += rel[2] * gp[0] - rel[0] * gp[2];
    sum_w[2] += rel[0] * gp[1] - rel[1] * gp[0];
    const float* q = &scene.rotations[i * 4];
    const double len = std::sqrt(double(q[0]) * q[0] + double(q[1]) * q[1] + double(q[2]) * q[2] + double(q[3]) * q[3]);
    const double w = q[0] / len, u[3] = {q[1] / len, q[2] / len, q[3] / len};
    const double gw = gp[8] * len, gu[3] = {gp[9] * len, gp[10] * len, gp[11] * len};
    const double ug[3] = {u[1] * gu[2] - u[2] * gu[1], u[2] * gu[0] - u[0] * gu[2], u[0] * gu[1] - u[1] * gu[0]};
    for (int k = 0; k < 3; ++k) sum_w[k] += 0.5 * (-gw * u[k] + w * gu[k] + ug[k]);
  }
  double expected[6], worst = 0.0, largest = 0.0;
  for (int r = 0; r < 3; ++r) {
    expected[r] = expected[3 + r] = 0.0;
    for (int c = 0; c < 3; ++c) {
      expected[r] += cam.rotation[r * 3 + c] * sum_d[c];
      expected[3 + r] += cam.rotation[r * 3 + c] * sum_w[c];
    }
  }
  std::printf("  pose gradient (lambda %.1f):", lambda);
  for (int k = 0; k < 6; ++k) {
    std::printf(" %c%c %.4g/%.4g", k < 3 ? 'd' : 'w', "xyz"[k % 3], g[k], expected[k]);
    worst = std::max(worst, std::fabs(g[k] - expected[k]));
    largest = std::max(largest, std::fabs(expected[k]));
  }
  std::printf("\n");
  EXPECT(largest > 0.0 && worst <= 1e-3 * largest, "pose gradient off by %.3g (largest %.3g)", worst, largest);
}

// One camera of a rendered pair moved 10 cm and about 1.7 degrees (3 m away:
// about 2 px at 128 px); with the gaussians frozen, refining that view's pose
// must bring it back.
void test_pose_recovery(gs::Device& dev) {
  gs::Dataset ds = test_dataset(128, 96, 2, false, 3);
  const gs::SplatScene scene = test_scene(300, 0, 9);
  render_targets(dev, ds, scene);
  const gs::Camera truth = gs::dataset_camera(ds, ds.frames[1]);
  gs::Dataset moved = ds;
  const double d[3] = {0.07, -0.05, 0.05}, rot[3] = {0.012, 0.02, -0.01};
  nudge_camera(moved, 1, d, rot);
  gs::TrainOptions o = frozen(0);
  o.iterations = 200;
  o.refine_poses = true;
  o.pose_from = 0;
  o.lr_pose_shift = 2e-3f;
  o.lr_pose_turn = 2e-3f;
  gs::Trainer t(dev, moved, scene, o);
  const float black[3] = {0, 0, 0};
  for (int s = 0; s < 150; ++s) t.step_on(1, black);
  double shift0, turn0, shift1, turn1;
  camera_error(gs::dataset_camera(moved, moved.frames[1]), truth, &shift0, &turn0);
  camera_error(t.camera_for(1), truth, &shift1, &turn1);
  std::printf("  pose recovery: %.1f cm, %.2f deg -> %.1f cm, %.2f deg (%u updates)\n", 100 * shift0, turn0,
              100 * shift1, turn1, t.stats().pose_updates);
  EXPECT(shift1 < 0.25 * shift0 && turn1 < 0.25 * turn0, "pose not recovered: %.3f m %.2f deg left", shift1, turn1);
  EXPECT(t.stats().pose_updates == 150, "pose updates %u", t.stats().pose_updates);
}

// Held-out views are aligned before evaluation: a held-out camera moved as
// above scores better after alignment than at its given pose.
void test_heldout_alignment(gs::Device& dev) {
  gs::Dataset ds = test_dataset(128, 96, 9, false, 5);  // frames 0 and 8 held out
  const gs::SplatScene scene = test_scene(300, 0, 9);
  render_targets(dev, ds, scene);
  const double d[3] = {0.03, -0.02, 0.02}, rot[3] = {0.004, 0.006, -0.003};
  nudge_camera(ds, 0, d, rot);
  gs::TrainOptions o = frozen(0);
  o.iterations = 100;
  o.refine_poses = true;
  o.eval_pose_steps = 12;
  gs::Trainer t(dev, ds, scene, o);
  const gs::EvalResult r = t.evaluate();
  std::printf("  held-out alignment: %u views, %.2f dB at the given poses -> %.2f dB aligned\n", r.views,
              r.psnr_unaligned, r.psnr);
  EXPECT(r.aligned && r.views == 2 && r.psnr > r.psnr_unaligned + 1.0, "alignment: %.2f -> %.2f dB",
         r.psnr_unaligned, r.psnr);
}

// Views rendered from a scene, one darkened and tinted as a camera's exposure
// and white balance would; with the gaussians frozen the trained views'
// gains come out in that ratio, their geometric mean held at 1.
void test_exposure_recovery(gs::Device& dev) {
  constexpr uint32_t w = 128, h = 96;
  gs::Dataset ds = test_dataset(w, h, 3, false, 3);  // frame 0 held out
  const gs::SplatScene scene = test_scene(300, 0, 9);
  render_targets(dev, ds, scene);
  const float dim[3] = {0.8f, 0.85f, 0.9f};
  for (size_t p = 0; p < size_t(w) * h; ++p) {
    for (int c = 0; c < 3; ++c) {
      uint8_t& v = ds.frames[1].rgba[p * 4 + c];
      v = uint8_t(std::lround(v * dim[c]));
    }
  }
  gs::TrainOptions o = frozen(0);
  o.iterations = 100;
  o.refine_exposure = true;
  o.exposure_from = 0;
  gs::Trainer t(dev, ds, scene, o);
  const float black[3] = {0, 0, 0};
  for (int s = 0; s < 30; ++s) t.step_on(1 + s % 2, black);
  float g1[3], g2[3], g0[3];
  t.view_gain(1, g1);
  t.view_gain(2, g2);
  t.view_gain(0, g0);  // held out, before every trained view: view 1's
  double worst = 0.0;
  for (int c = 0; c < 3; ++c) worst = std::max(worst, std::fabs(double(g1[c]) / g2[c] / dim[c] - 1.0));
  std::printf("  exposure: gains %.3f %.3f %.3f / %.3f %.3f %.3f, ratio off by %.2f%%, held-out %.3f, %u updates\n",
              g1[0], g1[1], g1[2], g2[0], g2[1], g2[2], 100.0 * worst, g0[1], t.stats().exposure_updates);
  EXPECT(worst < 0.02, "exposure gains off by %.2f%%", 100.0 * worst);
  EXPECT(std::fabs(std::log(double(g1[1])) + std::log(double(g2[1]))) < 1e-4, "gains not centred: %.4f %.4f", g1[1],
         g2[1]);
  EXPECT(g0[0] == g1[0] && g0[1] == g1[1] && g0[2] == g1[2], "held-out view takes its neighbour's gain");
}

}  // namespace

int main() {
  try {
    gs::Device dev(0);
    std::printf("device: %s\n", dev.name().c_str());
    std::fflush(stdout);
    test_pair_order(dev);
    test_frozen_step(dev);
    test_pose_gradient(dev, 0.0f);
    test_pose_gradient(dev, 0.2f);
    test_pose_recovery(dev);
    test_heldout_alignment(dev);
    test_exposure_recovery(dev);
    test_densify_at_capacity(dev);
    test_export_ply(dev);
    test_training_converges(dev, 1);
    test_training_converges(dev, 4);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_trainer: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
