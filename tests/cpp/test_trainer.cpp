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
#include "gs_hip.h"
#include "test_util.h"

namespace {

// Seeded test data in the shape of the reference's test generators
// (train_selftest.cpp:285-332): frames on an orbit of radius 3 about the
// origin whose height wobbles three times per turn, fx = fy = 0.9 W, photos of
// uniform random bytes (alpha 255 unless masked), every 8th frame held out
// when there are several; gaussians in a cube of half-width 0.7 with scales
// near 0.08, quaternions leaning to w, mid opacities and small SH.
gs::Dataset test_dataset(uint32_t w, uint32_t h, size_t frames, bool masks, uint32_t seed) {
  gs::Dataset ds;
  ds.name = "test";
  ds.width = w, ds.height = h;
  ds.fx = ds.fy = 0.9f * float(w);
  ds.cx = 0.5f * float(w), ds.cy = 0.5f * float(h);
  ds.has_masks = masks;
  ds.camera_extent = 3.3f;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<int> noise(0, 255);
  const float origin[3] = {0, 0, 0}, up[3] = {0, 1, 0};
  ds.frames.resize(frames);
  for (size_t i = 0; i < frames; ++i) {
    gs::DatasetFrame& f = ds.frames[i];
    const float turn = 6.2831853f * float(i) / float(frames);
    const float eye[3] = {3.0f * std::sin(turn), 0.6f + 0.3f * std::sin(3.0f * turn), 3.0f * std::cos(turn)};
    const gs::Camera c = gs::look_at(eye, origin, up, w, h, 60.0f);
    // Camera-to-world in OpenGL axes: the camera's x, -y, -z axes as columns, then its centre.
    for (int r = 0; r < 3; ++r) {
      const float col[4] = {c.rotation[r], -c.rotation[3 + r], -c.rotation[6 + r], eye[r]};
      std::copy_n(col, 4, &f.c2w[r * 4]);
    }
    const float bottom[4] = {0, 0, 0, 1};
    std::copy_n(bottom, 4, &f.c2w[12]);
    f.rgba.resize(size_t(w) * h * 4);
    std::generate(f.rgba.begin(), f.rgba.end(), [&] { return uint8_t(noise(gen)); });
    for (size_t p = 3; !masks && p < f.rgba.size(); p += 4) f.rgba[p] = 255;
    f.holdout = frames > 1 && i % 8 == 0;
  }
  return ds;
}

gs::SplatScene test_scene(uint32_t n, int degree, uint32_t seed) {
  gs::SplatScene s;
  s.convention = gs::AxisConvention::RUB;
  s.resize(n, degree);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> sym(-1.0f, 1.0f);
  const auto fill = [&](std::vector<float>& values, float scale, float offset) {
    for (float& v : values) v = offset + scale * sym(gen);
  };
  fill(s.positions, 0.7f, 0.0f);
  fill(s.log_scales, 0.06f, 0.08f);
  for (float& v : s.log_scales) v = std::log(v);
  fill(s.sh_dc, 0.8f, 0.0f);
  fill(s.opacity_logits, 0.6f, -0.2f);
  fill(s.rotations, 1.0f, 0.0f);
  for (uint32_t i = 0; i < n; ++i) s.rotations[size_t(i) * 4] += 1.5f;
  fill(s.sh_rest, 0.25f, 0.0f);
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

// Densification on every step of the run, for gaussians whose mean
// view-space gradient reaches `threshold` (0: all of them).
void densify_each_step(gs::TrainOptions& o, float threshold) {
  o.densify_grad_threshold = threshold;
  o.densify_interval = 1;
  o.densify_until = o.iterations + 100;
  o.densify_from = 0;
}

// Gaussians frozen, every view's pose refined from the first step on at
// `rate` for shifts and turns alike (0: the pose gradient is only measured).
gs::TrainOptions refining_poses(float rate) {
  gs::TrainOptions o = frozen(0);
  o.pose_from = 0;
  o.lr_pose_turn = rate;
  o.lr_pose_shift = rate;
  o.refine_poses = true;
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
  densify_each_step(o, 0.0f);
  o.lr_position = 1e-3f;
  o.max_gaussians = 60;  // 40 gaussians, every one cloned or split: more than fit
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 17), test_scene(40, 1, 19), o);
  const float bg[3] = {0, 0, 0};
  t.step_on(0, bg);
  t.step_on(0, bg);
  EXPECT(t.stats().slots == 60 && t.stats().gaussians == 60, "slots %u alive %u", t.stats().slots,
         t.stats().gaussians);
  EXPECT(std::isfinite(t.stats().loss), "loss finite");
  EXPECT(t.stats().densify_events >= 1, "densified");
}

// export_scene and export_ply (rows written on the device, compacted by a
// scan) against the raw slots read back with gs_hip_workspace_store: the
// alive slots in slot order, quaternions divided by their length, and
// export_ply byte for byte equal to save_ply of that scene in RDF; with
// pruned (free) slots among the alive ones (about half are pruned, none
// spawn) and SH degree 2.
void test_export_ply(gs::Device& dev) {
  gs::TrainOptions o = frozen(2);
  o.iterations = 4;
  o.max_gaussians = 400;
  densify_each_step(o, 1e9f);  // nothing qualifies: the densify passes only prune
  o.prune_opacity = 0.45f;
  o.lr_position = 1e-3f;
  o.lr_sh_rest = 1e-3f;
  gs::Trainer t(dev, test_dataset(64, 48, 1, false, 23), test_scene(150, 2, 29), o);
  const float bg[3] = {0, 0, 0};
  t.step_on(0, bg);
  t.step_on(0, bg);
  const uint32_t slots = t.stats().slots, K = t.sh_stride() / 3;
  std::vector<float> raw(size_t(slots) * 16), raw_sh(size_t(slots) * t.sh_stride());
  gs::check_status(gs_hip_workspace_store(t.workspace(), raw.data(), raw_sh.data(), slots));
  gs::SplatScene want;
  want.convention = gs::AxisConvention::RUB;
  std::vector<uint32_t> alive;
  for (uint32_t i = 0; i < slots; ++i)
    if (raw[size_t(i) * 16 + 7] > 0.5f) alive.push_back(i);
  want.resize(uint32_t(alive.size()), 2);
  for (size_t o2 = 0; o2 < alive.size(); ++o2) {
    const float* r = &raw[size_t(alive[o2]) * 16];
    const float norm = std::sqrt(r[8] * r[8] + r[9] * r[9] + r[10] * r[10] + r[11] * r[11]);
    for (int k = 0; k < 3; ++k) {
      want.positions[o2 * 3 + k] = r[k];
      want.log_scales[o2 * 3 + k] = r[4 + k];
      want.sh_dc[o2 * 3 + k] = r[12 + k];
    }
    for (int k = 0; k < 4; ++k) want.rotations[o2 * 4 + k] = r[8 + k] / norm;
    want.opacity_logits[o2] = r[3];
    for (uint32_t k = 0; k < 3 * K; ++k) want.sh_rest[o2 * 3 * K + k] = raw_sh[size_t(alive[o2]) * 3 * K + k];
  }
  const gs::SplatScene got_scene = t.export_scene();
  EXPECT(got_scene.count == want.count && got_scene.positions == want.positions &&
             got_scene.log_scales == want.log_scales && got_scene.rotations == want.rotations &&
             got_scene.opacity_logits == want.opacity_logits && got_scene.sh_dc == want.sh_dc &&
             got_scene.sh_rest == want.sh_rest,
         "export_scene differs from the stored slots (%u vs %u gaussians)", got_scene.count, want.count);
  gs::convert_convention(want, gs::AxisConvention::RDF);
  const std::vector<uint8_t> ply = t.export_ply(), ply_want = gs::save_ply(want);
  EXPECT(t.stats().pruned > 0 && want.count < slots, "free slots: %u alive of %u slots", want.count, slots);
  EXPECT(ply == ply_want, "export_ply: %zu bytes vs save_ply %zu bytes", ply.size(), ply_want.size());
  std::printf("  export: %u gaussians of %u slots, PLY %zu bytes, equal to the stored slots\n", want.count, slots,
              ply.size());
}

// A step of more views than there are training frames takes each training
// frame once: three training frames (frame 0 held out), eight views per step.
// The pose gradient is measured (rates 0) for each view of the step, so the
// pose updates count the views stepped.
void test_views_per_step_draws_distinct_frames(gs::Device& dev) {
  gs::TrainOptions o = refining_poses(0.0f);
  o.iterations = 12;
  o.views_per_step = 8;
  gs::Trainer t(dev, test_dataset(64, 48, 4, false, 41), test_scene(100, 0, 43), o);
  t.step();
  EXPECT(t.stats().step == 3 && t.stats().optimizer_steps == 1 && t.stats().pose_updates == 3,
         "one step over 3 training frames: %u views, %u optimizer steps, %u pose updates", t.stats().step,
         t.stats().optimizer_steps, t.stats().pose_updates);
  t.steps(3);
  EXPECT(t.stats().step == 12 && t.stats().pose_updates == 12, "four steps: %u views, %u pose updates",
         t.stats().step, t.stats().pose_updates);
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
  // The start: the right gaussians, each coordinate moved by up to 8 cm
  // (uniform, 4.6 cm rms), colourless and a quarter opaque.
  gs::SplatScene init = truth;
  std::minstd_rand jitter(53);
  std::uniform_real_distribution<float> shove(-0.08f, 0.08f);
  std::transform(init.positions.begin(), init.positions.end(), init.positions.begin(),
                 [&](float v) { return v + shove(jitter); });
  std::fill(init.sh_dc.begin(), init.sh_dc.end(), 0.0f);
  std::fill(init.sh_rest.begin(), init.sh_rest.end(), 0.0f);
  std::fill(init.opacity_logits.begin(), init.opacity_logits.end(), -1.1f);
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

// Small 3x3 algebra for the pose tests (row-major, double).
struct M3 {
  double a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double operator()(int r, int c) const { return a[r * 3 + c]; }
};

M3 mul(const M3& x, const M3& y) {
  M3 z;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) z.a[r * 3 + c] = x(r, 0) * y(0, c) + x(r, 1) * y(1, c) + x(r, 2) * y(2, c);
  return z;
}

M3 from_floats(const float* m) {
  M3 z;
  for (int k = 0; k < 9; ++k) z.a[k] = m[k];
  return z;
}

// The rotation by angle |w| about w, through its unit quaternion.
M3 turn_matrix(const double w[3]) {
  const double angle = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double s = angle > 0.0 ? std::sin(0.5 * angle) / angle : 0.5;
  const double q0 = std::cos(0.5 * angle), q1 = s * w[0], q2 = s * w[1], q3 = s * w[2];
  M3 z;
  const double m[9] = {1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3),     2 * (q1 * q3 + q0 * q2),
                       2 * (q1 * q2 + q0 * q3),     1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1),
                       2 * (q1 * q3 - q0 * q2),     2 * (q2 * q3 + q0 * q1),     1 - 2 * (q1 * q1 + q2 * q2)};
  std::copy_n(m, 9, z.a);
  return z;
}

// Frame f's camera moved as a pose correction moves it: world -> camera
// [R | t] becomes [E R | E t + d] with E = turn_matrix(w).  Written back as
// camera-to-world in OpenGL axes.
void nudge_camera(gs::Dataset& ds, size_t f, const double d[3], const double w[3]) {
  const gs::Camera cam = gs::dataset_camera(ds, ds.frames[f]);
  const M3 E = turn_matrix(w), R = mul(E, from_floats(cam.rotation));
  double t[3];
  for (int r = 0; r < 3; ++r) {
    t[r] = d[r] + E(r, 0) * cam.translation[0] + E(r, 1) * cam.translation[1] + E(r, 2) * cam.translation[2];
  }
  // camera -> world: rotation R^T (columns = camera axes; OpenGL flips y and z), centre -R^T t.
  const double flip[3] = {1.0, -1.0, -1.0};
  float* m = ds.frames[f].c2w;
  for (int r = 0; r < 3; ++r) {
    for (int axis = 0; axis < 3; ++axis) m[r * 4 + axis] = float(flip[axis] * R(axis, r));
    m[r * 4 + 3] = float(-(R(0, r) * t[0] + R(1, r) * t[1] + R(2, r) * t[2]));
  }
}

// Distance between camera centres, and the angle of the turn from one
// camera's orientation to the other's in degrees (|Ra - Rb|_F = 2 sqrt 2 sin(angle / 2)).
void camera_error(const gs::Camera& a, const gs::Camera& b, double* shift, double* turn) {
  double centre = 0.0, frob = 0.0;
  for (int k = 0; k < 3; ++k) centre += std::pow(double(a.position[k]) - b.position[k], 2);
  for (int k = 0; k < 9; ++k) frob += std::pow(double(a.rotation[k]) - b.rotation[k], 2);
  *shift = std::sqrt(centre);
  *turn = 2.0 * std::asin(std::min(1.0, std::sqrt(frob) / (2.0 * std::sqrt(2.0)))) * 180.0 / M_PI;
}

// The pose gradient (the reference's criterion, test_trainer.cpp:404-464)
// against the per-gaussian gradients: each of the six camera motions acts on
// the image as a rigid motion of every gaussian, so dL/d(motion) is the sum
// over gaussians of their gradients dotted with that motion's velocity.  For
// camera axis k (R the world -> camera rotation, c the camera centre, a =
// R^T e_k the axis in world coordinates): a shift d = e_k (x_cam = R p + t + d)
// is every position moving by a; a turn about k (x_cam = E (R p + t)) is every
// position moving by a x (p - c) and every quaternion by (0, a / 2) q.
// SH degree 0: colour does not depend on the view direction, which the pose
// gradient leaves out.
void test_pose_gradient(gs::Device& dev, float lambda) {
  gs::Dataset ds = test_dataset(64, 48, 2, true, 3);
  const gs::SplatScene scene = test_scene(120, 0, 9);
  const float bg[3] = {0.2f, 0.3f, 0.4f};
  gs::TrainOptions o = refining_poses(0.0f);
  o.lambda_dssim_end = lambda;
  o.lambda_dssim = lambda;
  o.masked = true;
  gs::Trainer t(dev, ds, scene, o);
  t.step_on(0, bg);
  const double* got = t.last_pose_gradient();
  const gs::Trainer::Gradients grads = t.first_step_gradients();
  const gs::Camera cam = gs::dataset_camera(ds, ds.frames[0]);
  double want[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    const double a[3] = {cam.rotation[k * 3 + 0], cam.rotation[k * 3 + 1], cam.rotation[k * 3 + 2]};
    for (uint32_t i = 0; i < scene.count; ++i) {
      const float* g = &grads.params[size_t(i) * 16];
      const float* p = &scene.positions[size_t(i) * 3];
      const float* q = &scene.rotations[size_t(i) * 4];
      const double rel[3] = {p[0] - double(cam.position[0]), p[1] - double(cam.position[1]),
                             p[2] - double(cam.position[2])};
      // velocity of the position under the turn: a x rel
      const double v[3] = {a[1] * rel[2] - a[2] * rel[1], a[2] * rel[0] - a[0] * rel[2], a[0] * rel[1] - a[1] * rel[0]};
      // velocity of the quaternion (w, x, y, z) under the turn: (0, a / 2) * q
      const double dq[4] = {-0.5 * (a[0] * q[1] + a[1] * q[2] + a[2] * q[3]),
                            0.5 * (a[0] * q[0] + a[1] * q[3] - a[2] * q[2]),
                            0.5 * (a[1] * q[0] + a[2] * q[1] - a[0] * q[3]),
                            0.5 * (a[2] * q[0] + a[0] * q[2] - a[1] * q[1])};
      for (int c = 0; c < 3; ++c) {
        want[k] += g[c] * a[c];
        want[3 + k] += g[c] * v[c];
      }
      for (int c = 0; c < 4; ++c) want[3 + k] += g[8 + c] * dq[c];
    }
  }
  double worst = 0.0, largest = 0.0;
  std::printf("  pose gradient (lambda %.1f):", lambda);
  for (int k = 0; k < 6; ++k) {
    std::printf(" %s%c %.4g/%.4g", k < 3 ? "shift " : "turn ", "xyz"[k % 3], got[k], want[k]);
    worst = std::max(worst, std::fabs(got[k] - want[k]));
    largest = std::max(largest, std::fabs(want[k]));
  }
  std::printf("\n");
  EXPECT(largest > 0.0 && worst <= 1e-3 * largest, "pose gradient off by %.3g (largest %.3g)", worst, largest);
}

// The second camera of a rendered pair moved about 9 cm and 1.5 degrees (3 m
// away: about 2 px at 128 px); with the gaussians frozen, refining that
// view's pose must take away at least three quarters of both errors.
void test_pose_recovery(gs::Device& dev) {
  gs::Dataset ds = test_dataset(128, 96, 2, false, 3);
  const gs::SplatScene scene = test_scene(300, 0, 9);
  render_targets(dev, ds, scene);
  const gs::Camera truth = gs::dataset_camera(ds, ds.frames[1]);
  const auto error = [&truth](const gs::Camera& c) {
    std::pair<double, double> e;
    camera_error(c, truth, &e.first, &e.second);
    return e;
  };
  const double shift[3] = {-0.06, 0.045, 0.055}, turn[3] = {-0.015, 0.01, 0.018};
  nudge_camera(ds, 1, shift, turn);
  const std::pair<double, double> before = error(gs::dataset_camera(ds, ds.frames[1]));
  constexpr uint32_t kSteps = 160;
  gs::TrainOptions o = refining_poses(2e-3f);
  o.iterations = kSteps + 80;
  gs::Trainer t(dev, ds, scene, o);
  const float black[3] = {0, 0, 0};
  for (uint32_t s = 0; s < kSteps; ++s) t.step_on(1, black);
  const std::pair<double, double> after = error(t.camera_for(1));
  std::printf("  pose recovery: %.1f cm, %.2f deg -> %.1f cm, %.2f deg (%u updates)\n", 100 * before.first,
              before.second, 100 * after.first, after.second, t.stats().pose_updates);
  EXPECT(after.first < 0.25 * before.first && after.second < 0.25 * before.second,
         "pose not recovered: %.3f m %.2f deg left", after.first, after.second);
  EXPECT(t.stats().pose_updates == kSteps, "pose updates %u", t.stats().pose_updates);
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
  const float tint[3] = {0.78f, 0.86f, 0.92f};  // view 1's exposure and white balance
  std::vector<uint8_t>& dimmed = ds.frames[1].rgba;
  for (size_t k = 0; k < dimmed.size(); ++k) {
    if (k % 4 != 3) dimmed[k] = uint8_t(std::lround(dimmed[k] * tint[k % 4]));
  }
  gs::TrainOptions o = frozen(0);
  o.exposure_from = 0;
  o.refine_exposure = true;
  o.iterations = 120;
  gs::Trainer t(dev, ds, scene, o);
  const float black[3] = {0, 0, 0};
  for (int s = 0; s < 36; ++s) t.step_on(s % 2 == 0 ? 2 : 1, black);
  // gain[v]: view v's gain; view 0 is held out and takes the gain of the
  // first trained view after it, view 1.
  float gain[3][3];
  for (int v = 0; v < 3; ++v) t.view_gain(v, gain[v]);
  double worst = 0.0;
  for (int c = 0; c < 3; ++c) worst = std::max(worst, std::fabs(double(gain[1][c]) / gain[2][c] / tint[c] - 1.0));
  std::printf("  exposure: gains %.3f %.3f %.3f / %.3f %.3f %.3f, ratio off by %.2f%%, held-out %.3f, %u updates\n",
              gain[1][0], gain[1][1], gain[1][2], gain[2][0], gain[2][1], gain[2][2], 100.0 * worst, gain[0][1],
              t.stats().exposure_updates);
  EXPECT(worst < 0.02, "exposure gains off by %.2f%%", 100.0 * worst);
  const double centre = std::log(double(gain[1][1])) + std::log(double(gain[2][1]));
  EXPECT(std::fabs(centre) < 1e-4, "gains not centred: %.4f %.4f", gain[1][1], gain[2][1]);
  EXPECT(std::equal(gain[0], gain[0] + 3, gain[1]), "held-out view takes its neighbour's gain");
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
    test_views_per_step_draws_distinct_frames(dev);
    test_training_converges(dev, 1);
    test_training_converges(dev, 4);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  if (test::failures() == 0) std::printf("test_trainer: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
