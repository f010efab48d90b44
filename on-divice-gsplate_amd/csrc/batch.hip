// Batches of scenes: several workspaces trained by one launch per stage.
//
// The batch is a launch context (a gs_hip_workspace that owns no model and no
// views) with tables large enough for up to 64 views of up to 32 scenes.  Each
// step lists the scenes' views in scene order (TableEntry), uploads their
// frames, buffers and models into the context's ViewDev / SceneDev tables
// and runs the same stage launches as a single workspace's step: per-slot
// stages take their scene from blockIdx.y, per-view stages their view, the
// rasterizers (view, tile) work items from one list over every view.  Pair
// counts stay on the device; an overflow anywhere leaves every model alone
// and the step runs again with the overflowing scenes' buffers grown.
#include <algorithm>
#include <vector>

#include "gs_internal.h"

struct gs_hip_batch {
  gs_hip_workspace ctx;
  std::vector<gs_hip_workspace*> scenes;
};

namespace {

int check_scene_frames(gs_hip_workspace* s, const gs_hip_frame* frames, uint32_t n) {
  if (n == 0 || n > s->views.size()) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: a scene's view count is 0 or exceeds its views");
  }
  for (uint32_t v = 0; v < n; ++v) {
    const gs_hip_frame& f = frames[v];
    if (f.slots > s->capacity || f.slots != frames[0].slots) {
      return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: slots differ within a scene or exceed its capacity");
    }
    if (f.cam.width != s->width || f.cam.height != s->height || f.tiles_x != s->tiles_x || f.tiles_y != s->tiles_y) {
      return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: a frame's image size differs from its scene");
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int gs_hip_batch_create(gs_hip_batch** out, gs_hip_workspace* const* scenes, uint32_t count) {
  gsh::clear_error();
  if (out == nullptr || scenes == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_create: null argument");
  *out = nullptr;
  if (count == 0 || count > gsh::kMaxBatchScenes) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_create: 1 to 32 scenes");
  }
  for (uint32_t s = 0; s < count; ++s) {
    if (scenes[s] == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_create: null scene");
    for (uint32_t t = 0; t < s; ++t) {
      if (scenes[t] == scenes[s]) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_create: a scene appears twice");
    }
  }
  auto* b = new gs_hip_batch();
  b->scenes.assign(scenes, scenes + count);
  gs_hip_workspace& c = b->ctx;
  c.borrowed_views = true;
  c.cus = scenes[0]->cus;
  c.tuning.raster_groups_per_cu = 0;
  c.tuning.depth_digit_bits = 8;
  c.tuning.tile_digit_bits = 8;
  c.tuning.graph = 0;
  c.tuning.bin_shift = 0;
  c.tuning.stored_state = 0;
  uint64_t tiles = 0;
  for (gs_hip_workspace* s : b->scenes) tiles += uint64_t(GS_HIP_MAX_VIEWS) * gsh::tile_count(s);
  int rc = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) == hipSuccess
               ? 0
               : gsh::set_error(GS_HIP_E_HIP, "hipStreamCreate failed");
  if (rc == 0 && tiles * 4u > 0xffffffffull) rc = gsh::set_error(GS_HIP_E_CAPACITY, "batch_create: too many tiles");
  if (rc == 0) rc = gsh::alloc_tables(&c, gsh::kMaxBatchViews, gsh::kMaxBatchScenes, uint32_t(tiles));
  if (rc == 0 && hipStreamSynchronize(c.stream) != hipSuccess) rc = gsh::set_error(GS_HIP_E_HIP, "batch init sync");
  if (rc != 0) {
    gs_hip_batch_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

int gs_hip_batch_destroy(gs_hip_batch* b) {
  if (b == nullptr) return 0;
  gs_hip_workspace& c = b->ctx;
  if (c.stream != nullptr) hipStreamSynchronize(c.stream);
  gsh::drop_graph(&c);
  gsh::release_tables(&c);
  gsh::release_timer(c.timer);
  if (c.stream != nullptr) hipStreamDestroy(c.stream);
  c.stream = nullptr;
  delete b;
  return 0;
}

int gs_hip_batch_train(gs_hip_batch* b, const gs_hip_frame* frames, const uint32_t* views, const void* const* targets,
                       uint64_t* pairs, double* losses) {
  if (b == nullptr || frames == nullptr || views == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: null argument");
  }
  gs_hip_workspace& c = b->ctx;
  uint32_t total = 0;
  for (size_t s = 0; s < b->scenes.size(); ++s) {
    const int rc = check_scene_frames(b->scenes[s], frames + total, views[s]);
    if (rc != 0) return rc;
    total += views[s];
    if (total > gsh::kMaxBatchViews) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: more than 64 views");
  }
  const uint32_t depth16 = frames[0].flags & GS_HIP_FLAG_DEPTH16;
  for (uint32_t i = 0; i < total; ++i) {
    if ((frames[i].flags & GS_HIP_FLAG_DEPTH16) != depth16) {
      return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_train: every frame must use the same depth key width");
    }
  }
  // The scenes' own streams may still run their loads or densification.
  for (gs_hip_workspace* s : b->scenes) GS_HIP_TRY(hipStreamSynchronize(s->stream));
  // Headroom over the largest count each scene has seen (trainer.cpp:588-592).
  for (gs_hip_workspace* s : b->scenes) {
    const uint64_t wanted = s->max_pairs_seen + s->max_pairs_seen / 8;
    if (wanted > gsh::view0(s).pair_capacity) {
      const int rc = gsh::reserve_pairs(s, wanted);
      if (rc != 0) return rc;
    }
  }
  std::vector<gsh::TableEntry> entries;
  entries.reserve(total);
  for (size_t s = 0, i = 0; s < b->scenes.size(); ++s) {
    for (uint32_t v = 0; v < views[s]; ++v, ++i) {
      const void* t = targets != nullptr ? targets[i] : nullptr;
      entries.push_back(gsh::TableEntry{b->scenes[s]->views[v], &frames[i], b->scenes[s], t});
    }
  }
  for (int attempt = 0;; ++attempt) {
    int rc = gsh::upload_entries(&c, entries);
    if (rc == 0) rc = gsh::launch_step(&c, total, depth16 != 0u, true, true);
    if (rc != 0) return rc;
    GS_HIP_TRY(hipStreamSynchronize(c.stream));
    bool overflow = false;
    std::vector<uint64_t> most(b->scenes.size(), 0);
    for (size_t s = 0, i = 0; s < b->scenes.size(); ++s) {
      for (uint32_t v = 0; v < views[s]; ++v, ++i) {
        gsh::View& view = *entries[i].view;
        view.pairs = c.host_totals[i * 16];
        most[s] = std::max(most[s], view.pairs);
        overflow = overflow || view.pairs > view.pair_capacity;
      }
      b->scenes[s]->max_pairs_seen = std::max(b->scenes[s]->max_pairs_seen, most[s]);
    }
    if (!overflow) break;
    if (attempt > 4) return gsh::set_error(GS_HIP_E_CAPACITY, "batch_train: pair buffers keep overflowing");
    for (size_t s = 0; s < b->scenes.size(); ++s) {
      gs_hip_workspace* sc = b->scenes[s];
      if (most[s] <= gsh::view0(sc).pair_capacity) continue;
      ++sc->overflow_steps;
      rc = gsh::reserve_pairs(sc, most[s]);
      if (rc != 0) return rc;
    }
  }
  // Each scene's views hold their step's totals and tile losses too, as after a step of their own.
  for (uint32_t i = 0; i < total; ++i) {
    const gsh::View& v = *entries[i].view;
    GS_HIP_TRY(hipMemcpyAsync(v.totals, c.totals_all + size_t(i) * 16, 64, hipMemcpyDeviceToDevice, c.stream));
    GS_HIP_TRY(hipMemcpyAsync(v.tile_loss, c.tile_loss_all + size_t(c.tile_base[i]) * 4, size_t(v.tile_grid) * 16,
                              hipMemcpyDeviceToDevice, c.stream));
  }
  GS_HIP_TRY(hipStreamSynchronize(c.stream));
  for (uint32_t i = 0; i < total; ++i) {
    if (pairs != nullptr) pairs[i] = entries[i].view->pairs;
    if (losses != nullptr) gsh::sum_losses(&c, i, losses + 3 * i);
  }
  return 0;
}

int gs_hip_batch_get_tuning(gs_hip_batch* b, gs_hip_tuning* out) {
  if (b == nullptr || out == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_get_tuning: null");
  *out = b->ctx.tuning;
  return 0;
}

int gs_hip_batch_set_tuning(gs_hip_batch* b, const gs_hip_tuning* t) {
  if (b == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_set_tuning: null batch");
  return gs_hip_workspace_set_tuning(&b->ctx, t);
}

int gs_hip_batch_timings(gs_hip_batch* b, int enable, const char** names, double* ms, int capacity, int* count) {
  if (b == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "batch_timings: null batch");
  return gs_hip_workspace_timings(&b->ctx, enable, names, ms, capacity, count);
}

}  // extern "C"
