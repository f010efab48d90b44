// Trains gaussians and writes a 3DGS .ply (the command surface of
// engine/cli/gs_train.cpp:1-345).
//
//   gs_train <dataset_dir> <out.ply> [--width 360] [--iters 3000] [--max-gaussians N] [--sh 3]
//            [--init-gaussians 150000] [--opacity-reset N (0 = off)] [--grad-threshold 2e-4]
//            [--dssim 0.2] [--dssim-end L] [--no-masks] [--sparse-adam] [--preview <dir>] [--eval-every 1000]
//            [--views-per-step V] [--raster-groups G] [--depth-bits B] [--tile-bits B] [--graph 0|1]
//            [--bin-shift S] [--stored-state 0|1] [--autotune]
//            [--refine-poses] [--pose-lr shift,turn] [--pose-lr-final F] [--pose-from N] [--eval-pose-steps N]
//            [--pose-out poses.json] [--refine-exposure] (a brightness gain per view and colour channel)
//
// The dataset is a nerfstudio transforms.json folder (gs::load_dataset); the
// initial gaussians come from its depth maps, sparse points or cameras.  The
// reference's --reduction / --alternatives (its backward variants) become the
// launch-shape flags.  With --synthetic instead of a dataset, the views are
// rendered on an orbit from a ground-truth scene:
//
//   gs_train --synthetic <out.ply> [--gt scene.ply|scene.spz | --gt-random N] [--init-random N]
//            [--views V] [--size WxH] [--iters I] [--sh D] [--seed S] [--dssim L] [--no-densify] [...]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <random>
#include <stdexcept>
#include <string>

#include "gs/dataset.h"
#include "gs/device.h"
#include "gs/renderer.h"
#include "gs/splat_scene.h"
#include "gs/trainer.h"

namespace {

gs::SplatScene random_scene(uint32_t n, int degree, uint32_t seed) {
  // Means uniform in the unit ball, log-scales N(-4, 0.5), quaternions
  // normalise(N(0,1)^4), logits N(0,1), DC U[0,1], SH rest N(0, 0.2).
  gs::SplatScene s;
  s.convention = gs::AxisConvention::RUB;
  s.resize(n, degree);
  std::mt19937 rng(seed);
  std::normal_distribution<float> nrm(0.0f, 1.0f);
  std::uniform_real_distribution<float> uni(0.0f, 1.0f);
  for (uint32_t i = 0; i < n; ++i) {
    float d[3] = {nrm(rng), nrm(rng), nrm(rng)};
    const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-12f;
    const float r = std::cbrt(uni(rng));
    for (int k = 0; k < 3; ++k) {
      s.positions[i * 3 + k] = d[k] / len * r;
      s.log_scales[i * 3 + k] = -4.0f + 0.5f * nrm(rng);
      s.sh_dc[i * 3 + k] = uni(rng);
    }
    float q[4] = {nrm(rng), nrm(rng), nrm(rng), nrm(rng)};
    const float ql = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f;
    for (int k = 0; k < 4; ++k) s.rotations[i * 4 + k] = q[k] / ql;
    s.opacity_logits[i] = nrm(rng);
  }
  for (float& v : s.sh_rest) v = 0.2f * nrm(rng);
  return s;
}

struct Args {
  bool synthetic = false;
  std::string dataset, out, gt_path, preview, pose_out;
  int pose_from = -1;
  uint32_t width = 360, iters = 3000, init_gaussians = 150000, eval_every = 1000;
  int reset_override = -1;
  float grad_threshold = -1.0f;
  gs::TrainOptions o;
  // synthetic views
  uint32_t gt_random = 200000, init_random = 0, views = 24, size_w = 640, size_h = 360, seed = 1;
  bool densify = true;
};

void usage(const char* argv0) {
  std::fprintf(stderr,
               "usage: %s <dataset_dir> <out.ply> [--width 360] [--iters 3000] [--max-gaussians N] [--sh 3]\n"
               "         [--init-gaussians N] [--opacity-reset N] [--grad-threshold G] [--dssim L] [--dssim-end L]\n"
               "         [--no-masks] [--sparse-adam] [--preview <dir>] [--eval-every N] [--views-per-step V]\n"
               "         [--raster-groups G] [--depth-bits B] [--tile-bits B] [--graph 0|1] [--bin-shift S]\n"
               "         [--stored-state 0|1] [--autotune]\n"
               "         [--refine-poses] [--pose-lr shift,turn] [--pose-lr-final F] [--pose-from N]\n"
               "         [--eval-pose-steps N] [--pose-out poses.json] [--refine-exposure]\n"
               "       %s --synthetic <out.ply> [--gt scene | --gt-random N] [--init-random N] [--views V]\n"
               "         [--size WxH] [--iters I] [--sh D] [--seed S] [--dssim L] [--no-densify] [...]\n",
               argv0, argv0);
}

Args parse(int argc, char** argv) {
  Args a;
  int i = 1;
  if (argc >= 3 && std::string(argv[1]) == "--synthetic") {
    a.synthetic = true;
    a.out = argv[2];
    i = 3;
    a.iters = 300;
    a.eval_every = 0;
    a.o.masked = false;
  } else if (argc >= 3 && argv[1][0] != '-') {
    a.dataset = argv[1];
    a.out = argv[2];
    i = 3;
  } else {
    throw std::invalid_argument("missing <dataset_dir> <out.ply>");
  }
  a.o.max_gaussians = 300000;
  for (; i < argc; ++i) {
    const std::string s = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) throw std::invalid_argument(s + " needs a value");
      return argv[++i];
    };
    if (s == "--width") a.width = uint32_t(std::atoi(next()));
    else if (s == "--iters") a.iters = uint32_t(std::atoi(next()));
    else if (s == "--max-gaussians") a.o.max_gaussians = uint32_t(std::atoi(next()));
    else if (s == "--sh") a.o.sh_degree = std::atoi(next());
    else if (s == "--init-gaussians") a.init_gaussians = uint32_t(std::atoi(next()));
    else if (s == "--opacity-reset") a.reset_override = std::atoi(next());
    else if (s == "--grad-threshold") a.grad_threshold = float(std::atof(next()));
    else if (s == "--dssim" || s == "--lambda") a.o.lambda_dssim = float(std::atof(next()));
    else if (s == "--dssim-end") a.o.lambda_dssim_end = float(std::atof(next()));
    else if (s == "--no-masks") a.o.masked = false;
    else if (s == "--sparse-adam") a.o.sparse_adam = true;
    else if (s == "--preview") a.preview = next();
    else if (s == "--eval-every") a.eval_every = uint32_t(std::atoi(next()));
    else if (s == "--views-per-step") a.o.views_per_step = uint32_t(std::atoi(next()));
    else if (s == "--raster-groups") a.o.raster_groups_per_cu = std::atoi(next());
    else if (s == "--depth-bits") a.o.depth_digit_bits = std::atoi(next());
    else if (s == "--tile-bits") a.o.tile_digit_bits = std::atoi(next());
    else if (s == "--graph") a.o.graph = std::atoi(next());
    else if (s == "--bin-shift") a.o.bin_shift = std::atoi(next());
    else if (s == "--stored-state") a.o.stored_state = std::atoi(next());
    else if (s == "--autotune") a.o.autotune = true;
    else if (s == "--refine-poses") a.o.refine_poses = true;
    else if (s == "--pose-lr") {
      if (std::sscanf(next(), "%f,%f", &a.o.lr_pose_shift, &a.o.lr_pose_turn) != 2) {
        throw std::invalid_argument("--pose-lr needs shift,turn");
      }
    } else if (s == "--pose-lr-final") a.o.lr_pose_final = float(std::atof(next()));
    else if (s == "--pose-from") a.pose_from = std::atoi(next());
    else if (s == "--eval-pose-steps") a.o.eval_pose_steps = uint32_t(std::atoi(next()));
    else if (s == "--pose-out") a.pose_out = next();
    else if (s == "--refine-exposure") a.o.refine_exposure = true;
    else if (a.synthetic && s == "--gt") a.gt_path = next();
    else if (a.synthetic && s == "--gt-random") a.gt_random = uint32_t(std::atoi(next()));
    else if (a.synthetic && s == "--init-random") a.init_random = uint32_t(std::atoi(next()));
    else if (a.synthetic && s == "--views") a.views = uint32_t(std::atoi(next()));
    else if (a.synthetic && s == "--size") {
      if (std::sscanf(next(), "%ux%u", &a.size_w, &a.size_h) != 2) throw std::invalid_argument("--size WxH");
    } else if (a.synthetic && s == "--seed") a.seed = uint32_t(std::atoi(next()));
    else if (a.synthetic && s == "--no-densify") a.densify = false;
    else throw std::invalid_argument("unknown option " + s);
  }
  return a;
}

// Keeps the schedule proportional for shorter or longer runs (the reference's scaled defaults).
void scale_schedule(gs::TrainOptions& o, uint32_t iters, int pose_from) {
  const gs::TrainOptions d = gs::scaled_options(iters);
  o.iterations = d.iterations;
  o.densify_from = d.densify_from;
  o.densify_until = d.densify_until;
  o.opacity_reset_interval = d.opacity_reset_interval;
  o.sh_degree_interval = d.sh_degree_interval;
  o.pose_from = pose_from >= 0 ? uint32_t(pose_from) : d.pose_from;
  o.exposure_from = d.exposure_from;
}

// Every frame's camera-to-world (OpenGL axes, row-major) after refinement, as
// the reference's --pose-out writes it (cli/gs_train.cpp:284-302).
void write_poses(const gs::Trainer& trainer, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (f == nullptr) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "{\"frames\": [");
  const gs::Dataset& d = trainer.dataset();
  for (size_t i = 0; i < d.frames.size(); ++i) {
    const gs::Camera c = trainer.camera_for(i);
    std::fprintf(f, "%s\n{\"file_path\": \"%s\", \"holdout\": %s, \"transform_matrix\": [", i ? "," : "",
                 d.frames[i].name.c_str(), d.frames[i].holdout ? "true" : "false");
    for (int r = 0; r < 3; ++r) {
      // columns W^T e0, -W^T e1, -W^T e2 and the centre
      std::fprintf(f, "%s[%.9g,%.9g,%.9g,%.9g]", r ? "," : "", c.rotation[r], -c.rotation[3 + r], -c.rotation[6 + r],
                   c.position[r]);
    }
    std::fprintf(f, ",[0,0,0,1]]}");
  }
  std::fprintf(f, "\n]}\n");
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  Args a;
  try {
    a = parse(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gs_train: %s\n", e.what());
    usage(argv[0]);
    return 2;
  }
  try {
    using clock = std::chrono::steady_clock;
    gs::Device device(0);
    std::printf("device: %s\n", device.name().c_str());
    gs::Dataset ds;
    gs::SplatScene init;
    if (!a.synthetic) {
      const auto t0 = clock::now();
      ds = gs::load_dataset(a.dataset, a.width);
      std::printf("dataset %s: %zu frames (%zu held out, %zu skipped), %ux%u, masks %d, depth %d, %zu points (%.1f s)\n",
                  ds.name.c_str(), ds.frames.size(), ds.holdout_count(), ds.skipped.size(), ds.width, ds.height,
                  int(ds.has_masks), int(ds.has_depth), ds.points.size() / 3,
                  std::chrono::duration<double>(clock::now() - t0).count());
      for (const std::string& s : ds.skipped) std::printf("  skipped %s\n", s.c_str());
      gs::InitOptions io;
      io.subject_only = a.o.masked;
      io.max_gaussians = a.init_gaussians;
      gs::InitStats st;
      init = gs::initialize_gaussians(ds, io, &st);
      std::printf("init from %s: %zu points -> %u gaussians (voxel %.4f)\n", st.source.c_str(), st.raw_points,
                  init.count, st.voxel_size);
      scale_schedule(a.o, a.iters, a.pose_from);
    } else {
      gs::SplatScene gt;
      if (!a.gt_path.empty()) {
        const auto bytes = gs::read_file(a.gt_path);
        gt = gs::load_splat_file(bytes.data(), bytes.size());
        gs::convert_convention(gt, gs::AxisConvention::RUB);
      } else {
        gt = random_scene(a.gt_random, a.o.sh_degree, 1001);
      }
      ds = gs::make_orbit_dataset(a.views, a.size_w, a.size_h, 3.0f, 0.5f, 60.0f, 8);
      {
        auto scene = gs::upload_scene(device, gt, a.size_w, a.size_h);
        gs::SplatRenderer r(device);
        r.attach(scene.get());
        const float black[3] = {0, 0, 0};
        for (gs::DatasetFrame& f : ds.frames) {
          f.rgba = r.render(gs::dataset_camera(ds, f), black);
          for (size_t p = 0; p < size_t(a.size_w) * a.size_h; ++p) f.rgba[p * 4 + 3] = 255;
        }
      }
      init = random_scene(a.init_random ? a.init_random : gt.count, a.o.sh_degree, 2002 + a.seed);
      scale_schedule(a.o, a.iters, a.pose_from);
      a.o.sh_degree_interval = 0;
      a.o.lambda_dssim_end = -1.0f;
      a.o.random_background = false;
      a.o.max_gaussians = std::max(a.o.max_gaussians, 2 * init.count);
      a.o.seed = a.seed;
      if (!a.densify) a.o.densify_from = a.o.densify_until = ~0u;
    }
    if (a.reset_override >= 0) a.o.opacity_reset_interval = uint32_t(a.reset_override);
    if (a.grad_threshold >= 0.0f) a.o.densify_grad_threshold = a.grad_threshold;
    const uint32_t width = ds.width, height = ds.height;
    gs::Trainer trainer(device, std::move(ds), init, a.o);
    const auto t_start = clock::now();
    uint32_t last_print = 0;
    while (!trainer.finished()) {
      const bool profile = trainer.stats().optimizer_steps % 100 == 50;
      trainer.step(profile);
      const gs::TrainStats& st = trainer.stats();
      if (profile) {
        const gs::PassTimes& p = trainer.last_profile();
        std::printf("  pass ms: project %.3f sort %.3f raster %.3f ssim %.3f backward %.3f optimize %.3f\n",
                    p.project, p.sort, p.rasterize, p.ssim, p.backward, p.optimize);
      }
      if (st.step / 100 != last_print / 100 || trainer.finished()) {
        std::printf("step %u: loss %.5f psnr %.2f ssim %.4f gaussians %u pairs %llu  %.2f ms/step\n", st.step,
                    st.loss, st.psnr, st.ssim, st.gaussians, (unsigned long long)st.pairs, st.step_ms);
      }
      if (a.eval_every > 0 && (st.step / a.eval_every != last_print / a.eval_every || trainer.finished())) {
        const gs::EvalResult e = trainer.evaluate();
        std::printf("  eval: %u views psnr %.2f ssim %.4f", e.views, e.psnr, e.ssim);
        if (e.aligned) std::printf(" (poses aligned; %.2f dB at the given poses)", e.psnr_unaligned);
        std::printf("\n");
        if (a.o.refine_poses) {
          std::printf("  poses: %u updates, shift mean %.2f max %.2f cm, turn mean %.3f max %.3f deg\n",
                      st.pose_updates, 100.0 * st.pose_shift_mean, 100.0 * st.pose_shift_max, st.pose_turn_mean,
                      st.pose_turn_max);
        }
        if (a.o.refine_exposure) {
          std::printf("  exposure: %u updates, brightness gains %.3f..%.3f, tint spread %.3f\n", st.exposure_updates,
                      st.exposure_low, st.exposure_high, st.exposure_tint);
        }
        if (!a.preview.empty()) {
          const gs::Dataset& d = trainer.dataset();
          const float black[3] = {0, 0, 0};
          for (size_t f = 0; f < d.frames.size(); f += std::max<size_t>(1, d.frames.size() / 4)) {
            char name[64];
            std::snprintf(name, sizeof(name), "/step%05u_view%03zu.ppm", st.step, f);
            gs::write_ppm(a.preview + name, trainer.render(trainer.camera_for(f), black), width, height);
          }
        }
      }
      last_print = st.step;
    }
    const gs::TrainStats& st = trainer.stats();
    const double secs = std::chrono::duration<double>(clock::now() - t_start).count();
    std::printf("trained %u views in %u optimizer steps, %.1f s (%.1f views/s); %u densify events, %u overflow reruns\n",
                st.step, st.optimizer_steps, secs, st.step / std::max(secs, 1e-9), st.densify_events,
                st.overflow_steps);
    if (!a.pose_out.empty()) write_poses(trainer, a.pose_out);
    const auto t_export = clock::now();
    const std::vector<uint8_t> ply = trainer.export_ply();  // rows written on the device
    gs::write_file(a.out, ply);
    std::printf("wrote %s: %u gaussians (%.0f ms)\n", a.out.c_str(), trainer.stats().gaussians,
                std::chrono::duration<double, std::milli>(clock::now() - t_export).count());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gs_train: %s\n", e.what());
    return 1;
  }
}
