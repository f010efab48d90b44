// gs_train <out.ply> [options]: trains gaussians against posed views and
// writes the result as a 3DGS PLY (the command surface of
// engine/cli/gs_train.cpp, with the views synthesised on an orbit: the
// transforms.json dataset loader is not part of this build).
//
//   --gt <scene.ply|scene.spz>  ground-truth scene the target views are rendered from
//   --gt-random N               or N seeded random gaussians (default 200000)
//   --init-random N             initial gaussians (default: as many as the ground truth)
//   --views V --size WxH --iters I --sh D --seed S --lambda L --no-densify --eval-every N
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <random>
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

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2 || argv[1][0] == '-') {
    std::fprintf(stderr,
                 "usage: %s <out.ply> [--gt scene.ply|--gt-random N] [--init-random N] [--views V] [--size WxH] "
                 "[--iters I] [--sh D] [--seed S] [--lambda L] [--no-densify] [--eval-every N]\n",
                 argv[0]);
    return 2;
  }
  try {
    std::string out = argv[1], gt_path;
    uint32_t gt_random = 200000, init_random = 0, views = 24, width = 640, height = 360, iters = 300, seed = 1;
    uint32_t eval_every = 0;
    int sh = 3;
    float lambda = 0.2f;
    bool densify = true;
    for (int i = 2; i < argc; ++i) {
      const std::string a = argv[i];
      auto next = [&]() -> const char* {
        if (i + 1 >= argc) throw std::invalid_argument("missing value after " + a);
        return argv[++i];
      };
      if (a == "--gt") gt_path = next();
      else if (a == "--gt-random") gt_random = uint32_t(std::atoi(next()));
      else if (a == "--init-random") init_random = uint32_t(std::atoi(next()));
      else if (a == "--views") views = uint32_t(std::atoi(next()));
      else if (a == "--size") {
        if (std::sscanf(next(), "%ux%u", &width, &height) != 2) throw std::invalid_argument("--size WxH");
      } else if (a == "--iters") iters = uint32_t(std::atoi(next()));
      else if (a == "--sh") sh = std::atoi(next());
      else if (a == "--seed") seed = uint32_t(std::atoi(next()));
      else if (a == "--lambda") lambda = float(std::atof(next()));
      else if (a == "--no-densify") densify = false;
      else if (a == "--eval-every") eval_every = uint32_t(std::atoi(next()));
      else throw std::invalid_argument("unknown option " + a);
    }
    gs::Device device(0);
    std::printf("device: %s\n", device.name().c_str());
    gs::SplatScene gt;
    if (!gt_path.empty()) {
      const auto bytes = gs::read_file(gt_path);
      gt = gs::load_splat_file(bytes.data(), bytes.size());
      gs::convert_convention(gt, gs::AxisConvention::RUB);
    } else {
      gt = random_scene(gt_random, sh, 1001);
    }
    gs::Dataset ds = gs::make_orbit_dataset(views, width, height, 3.0f, 0.5f, 60.0f, 8);
    {
      auto scene = gs::upload_scene(device, gt, width, height);
      gs::SplatRenderer r(device);
      r.attach(scene.get());
      const float black[3] = {0, 0, 0};
      for (gs::DatasetFrame& f : ds.frames) {
        f.rgba = r.render(gs::dataset_camera(ds, f), black);
        for (size_t p = 0; p < size_t(width) * height; ++p) f.rgba[p * 4 + 3] = 255;
      }
    }
    const gs::SplatScene init = random_scene(init_random ? init_random : gt.count, sh, 2002 + seed);
    gs::TrainOptions o = gs::scaled_options(iters);
    o.sh_degree = sh;
    o.sh_degree_interval = 0;
    o.lambda_dssim = lambda;
    o.lambda_dssim_end = -1.0f;
    o.random_background = false;
    o.max_gaussians = 2 * init.count;
    o.seed = seed;
    if (!densify) o.densify_from = o.densify_until = ~0u;
    gs::Trainer trainer(device, std::move(ds), init, o);
    while (!trainer.finished()) {
      const bool profile = trainer.stats().step % 100 == 50;
      trainer.step(profile);
      const gs::TrainStats& st = trainer.stats();
      if (profile) {
        const gs::PassTimes& p = trainer.last_profile();
        std::printf("  pass ms: project %.3f sort %.3f raster %.3f ssim %.3f backward %.3f optimize %.3f\n",
                    p.project, p.sort, p.rasterize, p.ssim, p.backward, p.optimize);
      }
      if (st.step % 100 == 0 || trainer.finished()) {
        std::printf("step %u: loss %.5f psnr %.2f ssim %.4f gaussians %u pairs %llu  %.2f ms/step\n", st.step,
                    st.loss, st.psnr, st.ssim, st.gaussians, (unsigned long long)st.pairs, st.step_ms);
      }
      if (eval_every > 0 && st.step % eval_every == 0) {
        const gs::EvalResult e = trainer.evaluate();
        std::printf("  eval: %u views psnr %.2f ssim %.4f\n", e.views, e.psnr, e.ssim);
      }
    }
    gs::SplatScene result = trainer.export_scene();
    gs::convert_convention(result, gs::AxisConvention::RDF);
    gs::write_file(out, gs::save_ply(result));
    std::printf("wrote %s: %u gaussians\n", out.c_str(), result.count);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gs_train: %s\n", e.what());
    return 1;
  }
}
