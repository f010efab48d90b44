// The trainer handle: gs::Trainer behind the C ABI, one call per method of
// the reference's training session (app/src/main/cpp/jni_bridge.cpp:815-1108,
// NativeTrainer.kt:8-47).  prepare loads the folder, builds the initial
// gaussians and the trainer exactly as the reference's prepare does
// (scaled_options for the run length, sparse Adam from 2000 steps, an
// initial budget of half the slots capped at 150000); preview renders a
// held-out view over the reference's dark preview background.  Exceptions
// from the C++ layer come back as status codes with their message.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "gs/dataset.h"
#include "gs/device.h"
#include "gs/splat_scene.h"
#include "gs/trainer.h"
#include "gs_hip.h"
#include "gs_internal.h"

struct gs_hip_trainer {
  std::unique_ptr<gs::Device> device;
  std::unique_ptr<gs::Trainer> trainer;
  std::vector<size_t> preview_frames;  // held-out frames, else frame 0
};

namespace {

// The C++ layer's exception categories back to status codes (the inverse of gs::check_status).
template <typename Body>
int guarded(Body body) {
  try {
    gsh::clear_error();
    body();
    return GS_HIP_OK;
  } catch (const gs::HipError& e) {
    return gsh::set_error(e.code, e.what());
  } catch (const std::invalid_argument& e) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, e.what());
  } catch (const std::logic_error& e) {
    return gsh::set_error(GS_HIP_E_STATE, e.what());
  } catch (const std::bad_alloc&) {
    return gsh::set_error(GS_HIP_E_CAPACITY, "out of host memory");
  } catch (const std::exception& e) {
    return gsh::set_error(GS_HIP_E_IO, e.what());
  }
}

void fill_stats(const gs::Trainer& t, gs_hip_trainer_stats* out) {
  if (out == nullptr) return;
  const gs::TrainStats& s = t.stats();
  *out = gs_hip_trainer_stats{};
  out->step = s.step;
  out->iterations = t.options().iterations;
  out->gaussians = s.gaussians;
  out->optimizer_steps = s.optimizer_steps;
  out->sh_degree = s.sh_degree;
  out->densify_events = s.densify_events;
  out->overflow_steps = s.overflow_steps;
  out->pairs = s.pairs;
  out->l1 = s.l1;
  out->psnr = s.psnr;
  out->ssim = s.ssim;
  out->loss = s.loss;
  out->step_ms = s.step_ms;
  out->finished = t.finished() ? 1 : 0;
}

}  // namespace

extern "C" {

int gs_hip_trainer_create(const char* dataset_dir, const gs_hip_trainer_options* options, gs_hip_trainer** out,
                          gs_hip_trainer_info* info) {
  if (out == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_create: out is NULL");
  *out = nullptr;
  if (dataset_dir == nullptr || options == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_create: dataset_dir and options are required");
  }
  const gs_hip_trainer_options& o = *options;
  if (o.iterations == 0 || o.max_gaussians == 0 || o.width == 0) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_create: iterations, max_gaussians and width must be positive");
  }
  if (o.sh_degree < 0 || o.sh_degree > 3) return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_create: sh_degree 0..3");
  if (o.views_per_step > GS_HIP_MAX_VIEWS) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_create: views_per_step above GS_HIP_MAX_VIEWS");
  }
  auto handle = std::make_unique<gs_hip_trainer>();
  const int rc = guarded([&] {
    // The folder first: a bad path fails before the device is touched.
    gs::Dataset ds = gs::load_dataset(dataset_dir, o.width);
    gs::TrainOptions t = gs::scaled_options(o.iterations);
    t.max_gaussians = o.max_gaussians;
    t.sh_degree = o.sh_degree;
    t.masked = o.masked != 0;
    t.refine_poses = o.refine_poses != 0;
    t.refine_exposure = o.refine_exposure != 0;
    t.sparse_adam = o.iterations >= 2000;
    t.views_per_step = std::max<uint32_t>(1, o.views_per_step);
    t.seed = o.seed;
    // The launch shape and backward timed on the first training views (the
    // session steps one view at a time by default, where the stored-state
    // backward fills the GPU better than the re-walk).
    t.autotune = true;
    gs::InitOptions init;
    init.max_gaussians = std::min<uint32_t>(t.max_gaussians / 2, 150000);
    init.subject_only = t.masked;
    init.seed = o.seed;
    gs::InitStats init_stats;
    const gs::SplatScene scene = gs::initialize_gaussians(ds, init, &init_stats);
    for (size_t i = 0; i < ds.frames.size(); ++i) {
      if (ds.frames[i].holdout) handle->preview_frames.push_back(i);
    }
    if (handle->preview_frames.empty()) handle->preview_frames.push_back(0);
    if (info != nullptr) {
      *info = gs_hip_trainer_info{};
      info->frames = uint32_t(ds.frames.size());
      info->holdout = uint32_t(ds.holdout_count());
      info->skipped = uint32_t(ds.skipped.size());
      info->width = ds.width;
      info->height = ds.height;
      info->masks = ds.has_masks ? 1 : 0;
      info->depth = ds.has_depth ? 1 : 0;
      info->points = uint32_t(ds.points.size() / 3);
      info->init_points = uint64_t(init_stats.raw_points);
      info->init_gaussians = scene.count;
      std::snprintf(info->init_source, sizeof(info->init_source), "%s", init_stats.source.c_str());
    }
    handle->device = std::make_unique<gs::Device>(o.device);
    handle->trainer = std::make_unique<gs::Trainer>(*handle->device, std::move(ds), scene, t);
  });
  if (rc == GS_HIP_OK) *out = handle.release();
  return rc;
}

int gs_hip_trainer_destroy(gs_hip_trainer* trainer) {
  if (trainer == nullptr) return GS_HIP_OK;
  const int rc = guarded([&] { trainer->trainer.reset(); });
  delete trainer;
  return rc;
}

int gs_hip_trainer_steps(gs_hip_trainer* trainer, uint32_t count, int profile, uint32_t* done,
                         gs_hip_trainer_stats* stats, gs_hip_pass_times* times) {
  if (trainer == nullptr || !trainer->trainer) return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_steps: no trainer");
  return guarded([&] {
    gs::Trainer& t = *trainer->trainer;
    uint32_t n = 0;
    if (profile != 0 && count > 0 && !t.finished()) {
      t.step(true);  // timed pass by pass, on its own
      n = 1;
    }
    n += t.steps(count - n);
    if (done != nullptr) *done = n;
    fill_stats(t, stats);
    if (times != nullptr) {
      const gs::PassTimes& p = t.last_profile();
      *times = gs_hip_pass_times{p.project, p.sort, p.rasterize, p.ssim, p.backward, p.optimize, p.valid ? 1 : 0};
    }
  });
}

int gs_hip_trainer_get_stats(gs_hip_trainer* trainer, gs_hip_trainer_stats* stats) {
  if (trainer == nullptr || !trainer->trainer || stats == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_get_stats: no trainer or no output");
  }
  return guarded([&] { fill_stats(*trainer->trainer, stats); });
}

int gs_hip_trainer_evaluate(gs_hip_trainer* trainer, gs_hip_trainer_eval* out) {
  if (trainer == nullptr || !trainer->trainer || out == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_evaluate: no trainer or no output");
  }
  return guarded([&] {
    const gs::EvalResult e = trainer->trainer->evaluate();
    *out = gs_hip_trainer_eval{e.psnr, e.ssim, e.l1, e.psnr_unaligned, e.views, e.aligned ? 1 : 0};
  });
}

int gs_hip_trainer_preview(gs_hip_trainer* trainer, uint32_t index, uint8_t* rgba, size_t bytes, uint32_t* width,
                           uint32_t* height) {
  if (trainer == nullptr || !trainer->trainer) return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_preview: no trainer");
  return guarded([&] {
    gs::Trainer& t = *trainer->trainer;
    const gs::Dataset& ds = t.dataset();
    if (width != nullptr) *width = ds.width;
    if (height != nullptr) *height = ds.height;
    if (rgba == nullptr) return;  // size query
    const size_t pixels = size_t(ds.width) * ds.height;
    if (bytes < 2 * pixels * 4) throw std::invalid_argument("trainer_preview: the buffer holds less than 2 W H 4 bytes");
    const size_t frame = trainer->preview_frames[index % trainer->preview_frames.size()];
    const float bg[3] = {0.08f, 0.09f, 0.11f};
    const std::vector<uint8_t> render = t.render(t.camera_for(frame), bg);
    std::memcpy(rgba, render.data(), pixels * 4);
    // The photo over the same background (through its mask when training masked).
    const uint8_t* photo = ds.frames[frame].rgba.data();
    uint8_t* target = rgba + pixels * 4;
    const bool masked = t.options().masked;
    for (size_t p = 0; p < pixels; ++p) {
      const float a = masked ? photo[p * 4 + 3] / 255.0f : 1.0f;
      for (int c = 0; c < 3; ++c) {
        target[p * 4 + c] = uint8_t(photo[p * 4 + c] * a + bg[c] * 255.0f * (1.0f - a) + 0.5f);
      }
      target[p * 4 + 3] = 255;
    }
  });
}

int gs_hip_trainer_export_ply(gs_hip_trainer* trainer, const char* path, uint32_t* count, uint64_t* bytes) {
  if (trainer == nullptr || !trainer->trainer || path == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "trainer_export_ply: no trainer or no path");
  }
  return guarded([&] {
    const std::vector<uint8_t> ply = trainer->trainer->export_ply();  // rows written on the device
    gs::write_file(path, ply);
    if (count != nullptr) *count = trainer->trainer->stats().gaussians;
    if (bytes != nullptr) *bytes = ply.size();
  });
}

}  // extern "C"
