// The launch-shape autotuner: times the batched step of the caller's views
// under candidate shapes and keeps the fastest.  Shapes: persistent
// rasterizer workgroups per CU (or one workgroup per tile), radix digit
// widths of the depth and tile sorts (or one counting pass), the pair bin
// size, the backward (re-walk or stored state) and graph replay of the whole
// step.  Raster tiles stay 16x16, the
// reference's (trainer.h:351); with bin_shift 0 the pairs and ranges are the
// reference's, with 1 or 2 each tile walks its 32x32 or 64x64 bin's list and
// draws the same splats in the same order.
//
// The search is coordinate-wise (one knob at a time, the others at their
// best so far), which needs a dozen shapes instead of the full product.  Each
// shape runs one untimed step (capture, caches) and then `repeats` timed
// steps with GS_HIP_FLAG_NO_UPDATE, so the model is not changed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gs_internal.h"

namespace {

double time_shape(gs_hip_workspace* ws, const gs_hip_tuning& t, const gs_hip_frame* frames, uint32_t count,
                  const void* const* targets, uint32_t repeats, int* rc) {
  *rc = gs_hip_workspace_set_tuning(ws, &t);
  if (*rc != 0) return 0.0;
  *rc = gsh::run_step(ws, frames, count, targets, true, false, nullptr, nullptr);
  if (*rc != 0) return 0.0;
  const auto start = std::chrono::steady_clock::now();
  for (uint32_t r = 0; r < repeats && *rc == 0; ++r) {
    *rc = gsh::run_step(ws, frames, count, targets, true, false, nullptr, nullptr);
  }
  const auto stop = std::chrono::steady_clock::now();
  return std::chrono::duration<double, std::milli>(stop - start).count() / double(std::max<uint32_t>(repeats, 1));
}

}  // namespace

extern "C" {

int gs_hip_autotune(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets,
                    uint32_t repeats, gs_hip_tuning* best, double* best_ms) {
  if (ws == nullptr || frames == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "autotune: null argument");
  if (count == 0 || count > ws->views.size()) return gsh::set_error(GS_HIP_E_ARGUMENT, "autotune: view count");
  std::vector<gs_hip_frame> probe(frames, frames + count);
  for (gs_hip_frame& f : probe) f.flags |= GS_HIP_FLAG_NO_UPDATE;
  repeats = std::max<uint32_t>(repeats, 1);
  gs_hip_tuning cur = ws->tuning;
  int rc = 0;
  double cur_ms = time_shape(ws, cur, probe.data(), count, targets, repeats, &rc);
  if (rc != 0) return rc;
  auto consider = [&](gs_hip_tuning cand) {
    if (rc != 0 || std::memcmp(&cand, &cur, sizeof(cand)) == 0) return;
    const double ms = time_shape(ws, cand, probe.data(), count, targets, repeats, &rc);
    if (rc == 0 && ms < cur_ms) {
      cur = cand;
      cur_ms = ms;
    }
  };
  // The bin size first: it changes how many pairs the sorts see, and so which
  // digit widths win.
  // The backward with it: the stored-state backward runs with 16 px bins only.
  for (uint32_t shift : {0u, 1u, 2u}) {
    gs_hip_tuning c = cur;
    c.bin_shift = shift;
    c.stored_state = 0u;
    consider(c);
  }
  {
    gs_hip_tuning c = cur;
    c.bin_shift = 0u;
    c.stored_state = 1u;
    consider(c);
  }
  for (uint32_t groups : {0u, 1u, 2u, 4u, 8u}) {
    gs_hip_tuning c = cur;
    c.raster_groups_per_cu = groups;
    consider(c);
  }
  for (uint32_t bits : {8u, 6u, 5u}) {
    gs_hip_tuning c = cur;
    c.depth_digit_bits = bits;
    consider(c);
  }
  for (uint32_t bits : {0u, 8u, 6u, 4u}) {
    gs_hip_tuning c = cur;
    c.tile_digit_bits = bits;
    consider(c);
  }
  for (uint32_t graph : {0u, 1u}) {
    gs_hip_tuning c = cur;
    c.graph = graph;
    consider(c);
  }
  if (rc != 0) return rc;
  rc = gs_hip_workspace_set_tuning(ws, &cur);
  if (rc != 0) return rc;
  if (best != nullptr) *best = cur;
  if (best_ms != nullptr) *best_ms = cur_ms;
  return 0;
}

}  // extern "C"
