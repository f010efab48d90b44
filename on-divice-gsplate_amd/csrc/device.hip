// Device, memory and workspace entry points of the C ABI; the device table
// the batched kernels read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <string>

#include "gs_common.h"
#include "gs_internal.h"

namespace gsh {

namespace {
thread_local int g_code = 0;
thread_local std::string g_message;
}  // namespace

int set_error(int code, const std::string& message) {
  g_code = code;
  g_message = message;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  return set_error(GS_HIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void clear_error() {
  g_code = 0;
  g_message.clear();
}

void timer_begin(gs_hip_workspace* ws, const char* name) {
  Timer& t = ws->timer;
  if (!t.enabled) return;
  if (t.used == t.start.size()) {
    // Timing-only events: without the system-scope fence, recording one
    // neither writes back and invalidates the caches nor opens a gap of
    // about 10 us before the next kernel (elapsed times are read after a
    // device synchronize).
    hipEvent_t a, b;
    if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess) {
      return;
    }
    t.start.push_back(a);
    t.stop.push_back(b);
    t.names.emplace_back();
  }
  t.names[t.used] = name;
  hipEventRecord(t.start[t.used], ws->stream);
}

void timer_end(gs_hip_workspace* ws) {
  Timer& t = ws->timer;
  if (!t.enabled || t.used >= t.stop.size()) return;
  hipEventRecord(t.stop[t.used], ws->stream);
  ++t.used;
}

void release_timer(Timer& t) {
  for (hipEvent_t e : t.start) hipEventDestroy(e);
  for (hipEvent_t e : t.stop) hipEventDestroy(e);
  t.start.clear();
  t.stop.clear();
  t.names.clear();
  t.used = 0;
  t.enabled = false;
}

void drop_graph(gs_hip_workspace* ws) {
  if (ws->graph_exec != nullptr) hipGraphExecDestroy(ws->graph_exec);
  if (ws->graph != nullptr) hipGraphDestroy(ws->graph);
  ws->graph_exec = nullptr;
  ws->graph = nullptr;
  ws->graph_signature = 0;
}

}  // namespace gsh

namespace {

// Fills `count` u32 with `value` (a kernel, not hipMemsetAsync, so the step
// stays capturable in a graph).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t count, uint32_t value) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

// Streaming copy, 16 B per thread and iteration.
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    dst[i] = src[i];
  }
}

template <typename T>
int alloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return 0;
}

template <typename T>
void release(T*& p) {
  if (p != nullptr) hipFree(p);
  p = nullptr;
}

}  // namespace

namespace gsh {

int fill_u32(gs_hip_workspace* ws, void* p, uint64_t words, uint32_t value) {
  if (words == 0) return 0;
  k_fill_u32<<<div_up(words, 256), 256, 0, ws->stream>>>(static_cast<uint32_t*>(p), words, value);
  GS_CHECK_LAUNCH("k_fill_u32");
  return 0;
}

int create_view(gs_hip_workspace* ws, uint32_t index, View** out) {
  *out = nullptr;
  auto* v = new View();
  const uint64_t cap = std::max<uint32_t>(ws->capacity, 1);
  const uint64_t pixels = uint64_t(ws->width) * ws->height;
  const uint64_t tiles = tile_count(ws);
  int rc = 0;
  auto A = [&](int r) {
    if (rc == 0) rc = r;
  };
  A(alloc(&v->records, cap * GS_HIP_RECORD_FLOATS));
  A(alloc(&v->footprint, cap * 8));
  A(alloc(&v->tiles, cap));
  A(alloc(&v->misc, cap));
  A(alloc(&v->depth_keys, cap));
  A(alloc(&v->depth_order, cap));
  A(alloc(&v->offsets, cap));
  A(alloc(&v->grad2d, cap * GS_HIP_GRAD2D_FLOATS));
  A(alloc(&v->flags, cap));
  A(alloc(&v->scan_tmp, cap + 1));
  A(alloc(&v->ranges, (tiles + 1) * 2));
  A(alloc(&v->image, pixels * 4));
  A(alloc(&v->last, pixels));
  A(alloc(&v->target, pixels));
  A(alloc(&v->dldc, pixels * 4));
  A(alloc(&v->ssim_maps, pixels * 9));
  A(alloc(&v->tile_fit, tiles * 8));
  A(alloc(&v->pose_grad, div_up(cap, 256) * 8));
  v->block_sums_entries = div_up(cap + 1, 4096) + 1024;
  A(alloc(&v->block_sums, v->block_sums_entries));
  v->totals = ws->totals_all + size_t(index) * 16;
  v->tile_loss = ws->tile_loss_all + size_t(index) * tiles * 4;
  v->capacity = ws->capacity;
  v->tile_grid = uint32_t(tiles);
  if (rc == 0) rc = reserve_sort(*v, std::max<uint64_t>(cap, ws->views.empty() ? 0 : view0(ws).pair_capacity));
  if (rc == 0 && !ws->views.empty() && view0(ws).pair_capacity > 0) {
    // Same pair capacity as the other views.
    const uint64_t n = view0(ws).pair_capacity;
    A(alloc(&v->pair_tiles, n));
    A(alloc(&v->pair_gaussians, n));
    if (rc == 0) v->pair_capacity = n;
  }
  if (rc == 0) rc = fill_u32(ws, v->target, pixels, 0u);
  if (rc == 0) rc = fill_u32(ws, v->totals, 16, 0u);
  if (rc == 0 && hipStreamSynchronize(ws->stream) != hipSuccess) rc = set_error(GS_HIP_E_HIP, "view init sync");
  if (rc != 0) {
    destroy_view(v);
    return rc;
  }
  *out = v;
  return 0;
}

void destroy_view(View* v) {
  if (v == nullptr) return;
  release(v->records);
  release(v->footprint);
  release(v->tiles);
  release(v->misc);
  release(v->depth_keys);
  release(v->depth_order);
  release(v->offsets);
  release(v->grad2d);
  release(v->grad_fx);
  release(v->bucket_state);
  release(v->flags);
  release(v->scan_tmp);
  release(v->pair_tiles);
  release(v->pair_gaussians);
  release(v->ranges);
  release(v->image);
  release(v->last);
  release(v->target);
  release(v->dldc);
  release(v->ssim_maps);
  release(v->tile_fit);
  release(v->pose_grad);
  release(v->block_sums);
  release(v->sort.keys_alt);
  release(v->sort.vals_alt);
  release(v->sort.hist);
  release(v->sort.digit_totals);
  release(v->sort.tile_hist);
  delete v;
}

int reserve_sort(View& v, uint64_t items) {
  SortScratch& s = v.sort;
  if (items <= s.items) return 0;
  release(s.keys_alt);
  release(s.vals_alt);
  release(s.hist);
  release(s.digit_totals);
  release(s.tile_hist);
  s = SortScratch{};
  const uint64_t blocks = (items + 4095) / 4096;
  const uint64_t cols = std::max<uint64_t>(div_up(items, kTileChunk), 1);
  int rc = alloc(&s.keys_alt, items);
  if (rc == 0) rc = alloc(&s.vals_alt, items);
  if (rc == 0) rc = alloc(&s.hist, std::max<uint64_t>(blocks, 1) * 256);
  if (rc == 0) rc = alloc(&s.digit_totals, 256);
  if (rc == 0 && v.tile_grid + 1 <= kCountingMaxBins) rc = alloc(&s.tile_hist, cols * (v.tile_grid + 1));
  if (rc != 0) return rc;
  s.tile_cols = s.tile_hist != nullptr ? uint32_t(cols) : 0u;
  s.items = items;
  return 0;
}

int reserve_pairs(gs_hip_workspace* ws, uint64_t pairs) {
  if (pairs <= view0(ws).pair_capacity) return 0;
  if (pairs >= (uint64_t(1) << 32)) return set_error(GS_HIP_E_CAPACITY, "reserve_pairs: more than 2^32 pairs");
  // Headroom: views differ in pair count (trainer.cpp:588-603 keeps 1/8 + 1/4).
  uint64_t n = std::min<uint64_t>(pairs + pairs / 4, (uint64_t(1) << 32) - 1);
  n = std::max<uint64_t>(n, 1 << 16);
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  drop_graph(ws);
  for (View* v : ws->views) {
    release(v->pair_tiles);
    release(v->pair_gaussians);
    v->pair_capacity = 0;
    int rc = alloc(&v->pair_tiles, n);
    if (rc == 0) rc = alloc(&v->pair_gaussians, n);
    if (rc == 0) rc = reserve_sort(*v, std::max<uint64_t>(n, ws->capacity));
    if (rc != 0) return rc;
    v->pair_capacity = n;
  }
  ws->table_signature = 0;  // the segments point at the old buffers
  return 0;
}

namespace {

uint32_t bins_along(uint32_t pixels, uint32_t shift) {
  const uint32_t side = uint32_t(GS_HIP_TILE) << shift;
  return (pixels + side - 1) / side;
}

ViewDev make_view_dev(const View& v, const gs_hip_frame& f, uint32_t* totals, float* tile_loss, const void* target,
                      bool counting_tile_sort, bool deterministic, uint32_t bin_shift, bool stored_state,
                      const uint32_t* step_overflow) {
  ViewDev d;
  std::memset(&d, 0, sizeof(d));
  d.f = f;
  d.records = reinterpret_cast<float4*>(v.records);
  d.footprint = reinterpret_cast<float4*>(v.footprint);
  d.tiles = v.tiles;
  d.misc = v.misc;
  d.depth_keys = v.depth_keys;
  d.depth_order = v.depth_order;
  d.offsets = v.offsets;
  d.grad2d = v.grad2d;
  d.grad_fx = deterministic ? v.grad_fx : nullptr;
  d.pair_tiles = v.pair_tiles;
  d.pair_gaussians = v.pair_gaussians;
  d.emit_tiles = counting_tile_sort ? v.sort.keys_alt : v.pair_tiles;
  d.emit_gaussians = counting_tile_sort ? v.sort.vals_alt : v.pair_gaussians;
  d.tile_hist = v.sort.tile_hist;
  d.tile_cols = v.sort.tile_cols;
  d.ranges = reinterpret_cast<uint2*>(v.ranges);
  d.image = reinterpret_cast<float4*>(v.image);
  d.last = v.last;
  d.target = target != nullptr ? static_cast<const uint32_t*>(target) : v.target;
  d.dldc = reinterpret_cast<float4*>(v.dldc);
  d.maps = v.ssim_maps;
  d.tile_loss = reinterpret_cast<float4*>(tile_loss);
  d.tile_fit = reinterpret_cast<float4*>(v.tile_fit);
  d.pose_grad = reinterpret_cast<float4*>(v.pose_grad);
  d.totals = totals;
  d.pair_capacity = v.pair_capacity;
  d.capacity = v.capacity;
  d.tile_count = v.tile_grid;
  d.bin_shift = bin_shift;
  d.bins_x = bins_along(f.cam.width, bin_shift);
  d.bin_count = d.bins_x * bins_along(f.cam.height, bin_shift);
  const float W = float(f.cam.width), H = float(f.cam.height);
  d.lim_x = 1.3f * std::max(f.cam.cx, W - f.cam.cx) / f.cam.fx;
  d.lim_y = 1.3f * std::max(f.cam.cy, H - f.cam.cy) / f.cam.fy;
  d.bucket_state = stored_state ? reinterpret_cast<float4*>(v.bucket_state) : nullptr;
  d.bucket_slots = stored_state ? v.bucket_slots : 0;
  d.step_overflow = step_overflow;
  return d;
}

uint64_t mix64(uint64_t h, uint64_t x) {
  h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  return h;
}

// Sort and scan segments of the entries (rebuilt only when their buffers move).
int upload_segments(gs_hip_workspace* ws, const std::vector<TableEntry>& entries) {
  const size_t n = entries.size();
  std::vector<SortSeg> depth(n), tile(n);
  std::vector<ScanSeg> scan(n);
  const size_t slots_offset = offsetof(ViewDev, f) + offsetof(gs_hip_frame, slots);
  for (size_t i = 0; i < n; ++i) {
    View& v = *entries[i].view;
    const uint32_t* slots = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(ws->dev_views + i) +
                                                              slots_offset);
    uint32_t* totals = ws->totals_all + i * 16;
    const uint32_t blocks = std::max<uint32_t>(div_up(v.sort.items, 4096), 1);
    depth[i] = SortSeg{v.depth_keys, v.depth_order, v.sort.keys_alt, v.sort.vals_alt, v.sort.hist,
                       v.sort.digit_totals, slots, v.capacity, blocks};
    tile[i] = SortSeg{v.pair_tiles, v.pair_gaussians, v.sort.keys_alt, v.sort.vals_alt, v.sort.hist,
                      v.sort.digit_totals, totals, uint32_t(std::min<uint64_t>(v.pair_capacity, 0xffffffffu)), blocks};
    // 16-bit depth keys carry the slot's tile count in their upper half (the sort ranks the
    // lower half only), so the scan reads the sorted keys in order instead of gathering.
    if ((entries[i].frame->flags & GS_HIP_FLAG_DEPTH16) != 0u) {
      scan[i] = ScanSeg{v.depth_keys, nullptr, v.offsets, v.block_sums, totals, slots, v.capacity,
                        v.block_sums_entries, 16u};
    } else {
      scan[i] = ScanSeg{v.tiles, v.depth_order, v.offsets, v.block_sums, totals, slots, v.capacity,
                        v.block_sums_entries, 0u};
    }
  }
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  GS_HIP_TRY(hipMemcpy(ws->dev_depth_sort, depth.data(), sizeof(SortSeg) * n, hipMemcpyHostToDevice));
  GS_HIP_TRY(hipMemcpy(ws->dev_tile_sort, tile.data(), sizeof(SortSeg) * n, hipMemcpyHostToDevice));
  GS_HIP_TRY(hipMemcpy(ws->dev_tile_scan, scan.data(), sizeof(ScanSeg) * n, hipMemcpyHostToDevice));
  return 0;
}

int bits_for(uint32_t n) {
  int bits = 0;
  while (bits < 32 && (uint64_t(1) << bits) < n) ++bits;
  return bits;
}

}  // namespace

int upload_entries(gs_hip_workspace* ws, const std::vector<TableEntry>& entries, int bin_shift) {
  const uint32_t count = uint32_t(entries.size());
  const uint32_t shift = bin_shift >= 0 ? uint32_t(bin_shift) : ws->tuning.bin_shift;
  if (count == 0 || count > ws->table_views) return set_error(GS_HIP_E_ARGUMENT, "view count out of range");
  // Shape, scenes (runs of entries of one scene) and each view's first tile in the per-tile arrays.
  StepShape shape;
  shape.views = count;
  std::vector<SceneDev> scenes;
  ws->tile_base.assign(count, 0);
  uint64_t sig = count;
  for (uint32_t i = 0; i < count; ++i) {
    const TableEntry& e = entries[i];
    const View& v = *e.view;
    if (i == 0 || entries[i - 1].scene != e.scene) {
      gs_hip_workspace* m = e.scene;
      SceneDev d;
      d.params = reinterpret_cast<float4*>(m->params);
      d.sh = m->sh;
      d.m_params = reinterpret_cast<float4*>(m->m_params);
      d.v_params = reinterpret_cast<float4*>(m->v_params);
      d.m_sh = m->m_sh;
      d.v_sh = m->v_sh;
      d.stats = reinterpret_cast<float4*>(m->stats);
      d.dir_grad = reinterpret_cast<float4*>(m->dir_grad);
      d.capacity = m->capacity;
      d.sh_coeffs = m->sh_coeffs;
      d.first_view = i;
      d.views = 0;
      scenes.push_back(d);
      shape.max_capacity = std::max(shape.max_capacity, m->capacity);
      shape.any_sh = shape.any_sh || m->sh_coeffs > 0;
    }
    ++scenes.back().views;
    if (scenes.back().views > GS_HIP_MAX_VIEWS) return set_error(GS_HIP_E_ARGUMENT, "more views of one scene than a step holds");
    ws->tile_base[i] = shape.total_tiles;
    shape.total_tiles += v.tile_grid;
    shape.work_items += quad_items(e.frame->tiles_x, e.frame->tiles_y);
    shape.max_quads = std::max(shape.max_quads, quad_items(e.frame->tiles_x, e.frame->tiles_y) / 4u);
    shape.max_tiles = std::max(shape.max_tiles, v.tile_grid);
    shape.max_bins = std::max(shape.max_bins, bins_along(e.frame->cam.width, shift) *
                                                  bins_along(e.frame->cam.height, shift));
    shape.max_pair_capacity = std::max(shape.max_pair_capacity, v.pair_capacity);
    sig = mix64(mix64(mix64(sig, reinterpret_cast<uintptr_t>(&v)), v.pair_capacity), v.sort.items);
    // The flags that change the launch sequence (a captured graph holds one).
    sig = mix64(sig, e.frame->flags & (GS_HIP_FLAG_DEPTH16 | GS_HIP_FLAG_POSE_GRAD | GS_HIP_FLAG_DETERMINISTIC));
    sig = mix64(sig, shift);
    shape.any_pose = shape.any_pose || (e.frame->flags & GS_HIP_FLAG_POSE_GRAD) != 0u;
    shape.deterministic = shape.deterministic || (e.frame->flags & GS_HIP_FLAG_DETERMINISTIC) != 0u;
  }
  shape.scenes = uint32_t(scenes.size());
  shape.tile_key_bits = uint32_t(bits_for(shape.max_bins + 1));
  shape.counting_tile_sort = ws->tuning.tile_digit_bits == 0u && shape.max_bins + 1 <= kCountingMaxBins;
  for (const TableEntry& e : entries) {
    shape.counting_tile_sort = shape.counting_tile_sort && e.view->sort.tile_hist != nullptr &&
                               uint64_t(e.view->sort.tile_cols) * kTileChunk >= e.view->pair_capacity;
  }
  if (shape.scenes > ws->table_scenes) return set_error(GS_HIP_E_ARGUMENT, "too many scenes in one step");
  if (shape.work_items > ws->work_entries || shape.total_tiles > ws->tile_loss_entries) {
    return set_error(GS_HIP_E_ARGUMENT, "the step's tiles exceed the launch context's tables");
  }
  if (shape.deterministic) {  // the fixed-point sums, made on first use
    for (const TableEntry& e : entries) {
      if (e.view->grad_fx == nullptr) {
        GS_HIP_TRY(hipStreamSynchronize(ws->stream));
        const uint64_t n = uint64_t(std::max<uint32_t>(e.view->capacity, 1)) * GS_HIP_GRAD2D_FLOATS;
        int rc = alloc(&e.view->grad_fx, n);
        if (rc == 0) rc = fill_u32(ws, reinterpret_cast<uint32_t*>(e.view->grad_fx), n * 2, 0u);
        if (rc != 0) return rc;
      }
    }
  }
  // Stored-state backward: each view's bucket states for its pair buffers
  // (4 KB per 64 pairs, per tile of a bin), and the launch context's bucket
  // list (a tile's first bucket has no stored state), made on first use and
  // grown with the pair buffers.
  // With coarse bins a tile's walk holds its bin's other tiles' splats, and
  // a bin's slots would be taken 4^bin_shift times over: those steps re-walk.
  shape.stored_state = ws->tuning.stored_state != 0u && !shape.deterministic && shift == 0u;
  if (shape.stored_state) {
    uint64_t listed = 0;
    for (const TableEntry& e : entries) {
      const uint32_t bins = bins_along(e.frame->cam.width, shift) * bins_along(e.frame->cam.height, shift);
      const uint64_t slots = bucket_slots_for(e.view->pair_capacity, bins, shift);
      if (e.view->bucket_slots < slots) {
        GS_HIP_TRY(hipStreamSynchronize(ws->stream));
        release(e.view->bucket_state);
        e.view->bucket_slots = 0;
        const int rc = alloc(&e.view->bucket_state, slots * kBucketPixels * 4);
        if (rc != 0) return rc;
        e.view->bucket_slots = slots;
      }
      listed += slots + e.view->tile_grid;
      sig = mix64(mix64(sig, reinterpret_cast<uintptr_t>(e.view->bucket_state)), e.view->bucket_slots);
    }
    if (ws->bucket_list_entries < listed) {
      GS_HIP_TRY(hipStreamSynchronize(ws->stream));
      release(ws->bucket_list);
      ws->bucket_list_entries = 0;
      const int rc = alloc(&ws->bucket_list, listed);
      if (rc != 0) return rc;
      ws->bucket_list_entries = listed;
    }
    sig = mix64(mix64(sig, reinterpret_cast<uintptr_t>(ws->bucket_list)), ws->bucket_list_entries);
  }
  if (sig != ws->table_signature || ws->table_signature == 0) {
    const int rc = upload_segments(ws, entries);
    if (rc != 0) return rc;
    ws->table_signature = sig == 0 ? 1 : sig;
    drop_graph(ws);
  }
  ws->shape = shape;
  // A staging slot is reused only after its previous copy has run.
  const uint32_t s = ws->staging_next;
  ws->staging_next = (s + 1) % kStagingSlots;
  GS_HIP_TRY(hipEventSynchronize(ws->staging_done[s]));
  ViewDev* host = ws->host_views + size_t(s) * ws->table_views;
  for (uint32_t i = 0; i < count; ++i) {
    const TableEntry& e = entries[i];
    host[i] = make_view_dev(*e.view, *e.frame, ws->totals_all + size_t(i) * 16,
                            ws->tile_loss_all + size_t(ws->tile_base[i]) * 4, e.target, shape.counting_tile_sort,
                            shape.deterministic, shift, shape.stored_state,
                            ws->work_counters + kStepOverflowCounter);
  }
  SceneDev* hs = ws->host_scenes + size_t(s) * ws->table_scenes;
  std::copy(scenes.begin(), scenes.end(), hs);
  GS_HIP_TRY(hipMemcpyAsync(ws->dev_views, host, sizeof(ViewDev) * count, hipMemcpyHostToDevice, ws->stream));
  GS_HIP_TRY(hipMemcpyAsync(ws->dev_scenes, hs, sizeof(SceneDev) * scenes.size(), hipMemcpyHostToDevice, ws->stream));
  GS_HIP_TRY(hipEventRecord(ws->staging_done[s], ws->stream));
  return 0;
}

int upload_table(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets,
                 int bin_shift) {
  if (count == 0 || count > ws->views.size()) return set_error(GS_HIP_E_ARGUMENT, "view count out of range");
  std::vector<TableEntry> entries(count);
  for (uint32_t v = 0; v < count; ++v) {
    entries[v] = TableEntry{ws->views[v], &frames[v], ws, targets != nullptr ? targets[v] : nullptr};
  }
  return upload_entries(ws, entries, bin_shift);
}

int alloc_tables(gs_hip_workspace* ws, uint32_t views, uint32_t scenes, uint32_t tile_entries) {
  ws->table_views = views;
  ws->table_scenes = scenes;
  // Quads of 2x2 tiles: at most 4x the tiles (a grid of one tile).
  ws->work_entries = tile_entries * 4u;
  ws->tile_loss_entries = tile_entries;
  int rc = 0;
  auto A = [&](int r) {
    if (rc == 0) rc = r;
  };
  A(alloc(&ws->dev_views, views));
  A(alloc(&ws->dev_scenes, scenes));
  A(alloc(&ws->dev_depth_sort, views));
  A(alloc(&ws->dev_tile_sort, views));
  A(alloc(&ws->dev_tile_scan, views));
  A(alloc(&ws->dev_scan_tmp, 1));
  A(alloc(&ws->work, std::max<uint32_t>(ws->work_entries, 1)));
  A(alloc(&ws->work_counters, kWorkCounters));
  A(alloc(&ws->totals_all, size_t(16) * views));
  A(alloc(&ws->tile_loss_all, size_t(std::max<uint32_t>(tile_entries, 1)) * 4));
  auto pinned = [&](void** p, size_t bytes) {
    if (rc == 0 && hipHostMalloc(p, bytes) != hipSuccess) rc = set_error(GS_HIP_E_HIP, "hipHostMalloc failed");
  };
  pinned(reinterpret_cast<void**>(&ws->host_views), sizeof(ViewDev) * views * kStagingSlots);
  pinned(reinterpret_cast<void**>(&ws->host_scenes), sizeof(SceneDev) * scenes * kStagingSlots);
  pinned(reinterpret_cast<void**>(&ws->host_totals), size_t(64) * views);
  pinned(reinterpret_cast<void**>(&ws->host_tile_loss), size_t(std::max<uint32_t>(tile_entries, 1)) * 16);
  for (uint32_t s = 0; s < kStagingSlots && rc == 0; ++s) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      rc = set_error(GS_HIP_E_HIP, "hipEventCreate failed");
    } else {
      ws->staging_done.push_back(e);
    }
  }
  if (rc == 0) rc = fill_u32(ws, ws->totals_all, size_t(16) * views, 0);
  if (rc == 0) rc = fill_u32(ws, ws->work_counters, kWorkCounters, 0);
  return rc;
}

void release_tables(gs_hip_workspace* ws) {
  release(ws->dev_views);
  release(ws->dev_scenes);
  release(ws->dev_depth_sort);
  release(ws->dev_tile_sort);
  release(ws->dev_tile_scan);
  release(ws->dev_scan_tmp);
  release(ws->work);
  release(ws->work_counters);
  release(ws->bucket_list);
  ws->bucket_list_entries = 0;
  release(ws->totals_all);
  release(ws->tile_loss_all);
  if (ws->host_views != nullptr) hipHostFree(ws->host_views);
  if (ws->host_scenes != nullptr) hipHostFree(ws->host_scenes);
  if (ws->host_totals != nullptr) hipHostFree(ws->host_totals);
  if (ws->host_tile_loss != nullptr) hipHostFree(ws->host_tile_loss);
  ws->host_views = nullptr;
  ws->host_scenes = nullptr;
  ws->host_totals = nullptr;
  ws->host_tile_loss = nullptr;
  for (hipEvent_t e : ws->staging_done) hipEventDestroy(e);
  ws->staging_done.clear();
  for (hipEvent_t e : ws->timer.start) hipEventDestroy(e);
  for (hipEvent_t e : ws->timer.stop) hipEventDestroy(e);
  ws->timer.start.clear();
  ws->timer.stop.clear();
}

}  // namespace gsh

extern "C" {

int gs_hip_init(int device) {
  gsh::clear_error();
  int count = 0;
  GS_HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return gsh::set_error(GS_HIP_E_ARGUMENT, "gs_hip_init: no such device");
  GS_HIP_TRY(hipSetDevice(device));
  return 0;
}

int gs_hip_device_name(char* buffer, size_t size) {
  int dev = 0;
  GS_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GS_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (buffer != nullptr && size > 0) {
    std::snprintf(buffer, size, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  return 0;
}

int gs_hip_malloc(void** ptr, size_t bytes) {
  if (ptr == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "gs_hip_malloc: null ptr");
  GS_HIP_TRY(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return 0;
}

int gs_hip_free(void* ptr) {
  if (ptr != nullptr) GS_HIP_TRY(hipFree(ptr));
  return 0;
}

// The plain copies wait for all device work first: workspaces run on their own
// non-blocking streams, which a null-stream copy would not order against.
int gs_hip_upload(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_download(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int gs_hip_memset(void* dst, int value, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemset(dst, value, bytes));
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_copy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  if (dst == nullptr || src == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "gs_hip_copy: null pointer");
  GS_HIP_TRY(hipDeviceSynchronize());
  const bool vec = (reinterpret_cast<uintptr_t>(dst) % 16 == 0) && (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                   bytes % 16 == 0;
  if (vec) {
    const uint64_t n = bytes / 16;
    const uint32_t blocks = uint32_t(std::min<uint64_t>(gsh::div_up(n, 256), 65536));
    k_copy16<<<blocks, 256>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    GS_CHECK_LAUNCH("k_copy16");
  } else {
    GS_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  }
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_synchronize(void) {
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_last_error(char* buffer, size_t size) {
  if (buffer != nullptr && size > 0) std::snprintf(buffer, size, "%s", gsh::g_message.c_str());
  return gsh::g_code;
}

int gs_hip_workspace_create(gs_hip_workspace** out, uint32_t capacity, int sh_degree, uint32_t width,
                            uint32_t height) {
  gsh::clear_error();
  if (out == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: null handle");
  *out = nullptr;
  if (sh_degree < 0 || sh_degree > 3) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: sh_degree 0..3");
  if (width == 0 || height == 0 || width > 32768 || height > 32768) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: image size out of range");
  }
  if (capacity > (1u << 30)) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_create: capacity too large");
  auto* ws = new gs_hip_workspace();
  ws->capacity = capacity;
  ws->sh_degree = sh_degree;
  ws->sh_coeffs = sh_degree == 0 ? 0u : uint32_t((sh_degree + 1) * (sh_degree + 1) - 1);
  ws->width = width;
  ws->height = height;
  ws->tiles_x = gsh::div_up(width, GS_HIP_TILE);
  ws->tiles_y = gsh::div_up(height, GS_HIP_TILE);
  ws->tuning.raster_groups_per_cu = 0;
  ws->tuning.depth_digit_bits = 8;
  ws->tuning.tile_digit_bits = 8;
  ws->tuning.graph = 0;
  ws->tuning.bin_shift = 0;
  ws->tuning.stored_state = 0;
  const uint64_t cap = std::max<uint32_t>(capacity, 1);
  const uint64_t shn = cap * ws->sh_coeffs * 3;
  const uint64_t tiles = gsh::tile_count(ws);
  int rc = 0;
  auto A = [&](int r) {
    if (rc == 0) rc = r;
  };
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    ws->cus = uint32_t(prop.multiProcessorCount);
  }
  if (ws->cus == 0) ws->cus = 256;
  A(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking) == hipSuccess
        ? 0
        : gsh::set_error(GS_HIP_E_HIP, "hipStreamCreate failed"));
  A(alloc(&ws->params, cap * 16));
  A(alloc(&ws->sh, shn));
  A(alloc(&ws->m_params, cap * 16));
  A(alloc(&ws->v_params, cap * 16));
  A(alloc(&ws->m_sh, shn));
  A(alloc(&ws->v_sh, shn));
  A(alloc(&ws->stats, cap * 4));
  if (ws->sh_coeffs > 0) A(alloc(&ws->dir_grad, cap * uint64_t(gsh::kShGroups) * 4));
  if (tiles > gsh::kMaxTilesPerView) A(gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: more than 65535 tiles"));
  A(gsh::alloc_tables(ws, GS_HIP_MAX_VIEWS, 1, uint32_t(tiles) * GS_HIP_MAX_VIEWS));
  if (rc == 0) {
    for (void* p : {(void*)ws->params, (void*)ws->m_params, (void*)ws->v_params}) A(gsh::fill_u32(ws, p, cap * 16, 0));
    for (void* p : {(void*)ws->sh, (void*)ws->m_sh, (void*)ws->v_sh}) A(gsh::fill_u32(ws, p, shn, 0));
    A(gsh::fill_u32(ws, ws->stats, cap * 4, 0));
  }
  gsh::View* v = nullptr;
  if (rc == 0) rc = gsh::create_view(ws, 0, &v);
  if (v != nullptr) ws->views.push_back(v);
  // First guess before any count: a few tiles per gaussian (trainer.cpp:816-818).
  if (rc == 0) rc = gsh::reserve_pairs(ws, std::max<uint64_t>(uint64_t(capacity) * 4, 1 << 16));
  if (rc == 0 && hipStreamSynchronize(ws->stream) != hipSuccess) A(gsh::set_error(GS_HIP_E_HIP, "workspace init sync"));
  if (rc != 0) {
    gs_hip_workspace_destroy(ws);
    return rc;
  }
  *out = ws;
  return 0;
}

int gs_hip_workspace_views(gs_hip_workspace* ws, uint32_t count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_views: null workspace");
  if (count == 0 || count > GS_HIP_MAX_VIEWS) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_views: 1 to " + std::to_string(GS_HIP_MAX_VIEWS) + " views");
  }
  while (ws->views.size() < count) {
    gsh::View* v = nullptr;
    const int rc = gsh::create_view(ws, uint32_t(ws->views.size()), &v);
    if (rc != 0) return rc;
    ws->views.push_back(v);
    ws->table_signature = 0;
    gsh::drop_graph(ws);
  }
  return 0;
}

int gs_hip_workspace_destroy(gs_hip_workspace* ws) {
  if (ws == nullptr) return 0;
  hipDeviceSynchronize();
  gsh::drop_graph(ws);
  if (!ws->borrowed_views)
    for (gsh::View* v : ws->views) gsh::destroy_view(v);
  ws->views.clear();
  release(ws->params);
  release(ws->sh);
  release(ws->m_params);
  release(ws->v_params);
  release(ws->m_sh);
  release(ws->v_sh);
  release(ws->stats);
  release(ws->dir_grad);
  gsh::release_tables(ws);
  gsh::release_timer(ws->timer);
  if (ws->stream != nullptr) hipStreamDestroy(ws->stream);
  delete ws;
  return 0;
}

int gs_hip_workspace_view_buffer(gs_hip_workspace* ws, uint32_t view, int which, void** ptr, size_t* bytes) {
  if (ws == nullptr || ptr == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: null");
  if (view >= ws->views.size()) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: no such view");
  const gsh::View& v = *ws->views[view];
  const size_t cap = ws->capacity, pixels = size_t(ws->width) * ws->height;
  const size_t tiles = gsh::tile_count(ws), shn = cap * ws->sh_coeffs * 3;
  void* p = nullptr;
  size_t b = 0;
  switch (which) {
    case GS_HIP_BUF_PARAMS: p = ws->params, b = cap * 64; break;
    case GS_HIP_BUF_SH: p = ws->sh, b = shn * 4; break;
    case GS_HIP_BUF_M_PARAMS: p = ws->m_params, b = cap * 64; break;
    case GS_HIP_BUF_V_PARAMS: p = ws->v_params, b = cap * 64; break;
    case GS_HIP_BUF_M_SH: p = ws->m_sh, b = shn * 4; break;
    case GS_HIP_BUF_V_SH: p = ws->v_sh, b = shn * 4; break;
    case GS_HIP_BUF_STATS: p = ws->stats, b = cap * 16; break;
    case GS_HIP_BUF_RECORDS: p = v.records, b = cap * GS_HIP_RECORD_FLOATS * 4; break;
    case GS_HIP_BUF_FOOTPRINT: p = v.footprint, b = cap * 32; break;
    case GS_HIP_BUF_TILES: p = v.tiles, b = cap * 4; break;
    case GS_HIP_BUF_DEPTH_KEYS: p = v.depth_keys, b = cap * 4; break;
    case GS_HIP_BUF_DEPTH_ORDER: p = v.depth_order, b = cap * 4; break;
    case GS_HIP_BUF_OFFSETS: p = v.offsets, b = cap * 4; break;
    case GS_HIP_BUF_GRAD2D: p = v.grad2d, b = cap * GS_HIP_GRAD2D_FLOATS * 4; break;
    case GS_HIP_BUF_PAIR_TILES: p = v.pair_tiles, b = v.pair_capacity * 4; break;
    case GS_HIP_BUF_PAIR_GAUSSIANS: p = v.pair_gaussians, b = v.pair_capacity * 4; break;
    case GS_HIP_BUF_RANGES: p = v.ranges, b = (tiles + 1) * 8; break;
    case GS_HIP_BUF_IMAGE: p = v.image, b = pixels * 16; break;
    case GS_HIP_BUF_LAST: p = v.last, b = pixels * 4; break;
    case GS_HIP_BUF_TARGET: p = v.target, b = pixels * 4; break;
    case GS_HIP_BUF_DLDC: p = v.dldc, b = pixels * 16; break;
    case GS_HIP_BUF_TILE_LOSS: p = v.tile_loss, b = tiles * 16; break;
    case GS_HIP_BUF_TOTALS: p = v.totals, b = 64; break;
    case GS_HIP_BUF_TILE_ORDER:
      p = ws->work, b = size_t(gsh::quad_items(ws->tiles_x, ws->tiles_y)) * 4 * ws->views.size();
      break;
    case GS_HIP_BUF_TILE_FIT: p = v.tile_fit, b = tiles * 32; break;
    case GS_HIP_BUF_POSE_GRAD: p = v.pose_grad, b = gsh::div_up(uint32_t(std::max<size_t>(cap, 1)), 256) * 32; break;
    default: return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: unknown buffer");
  }
  *ptr = p;
  if (bytes != nullptr) *bytes = b;
  return 0;
}

int gs_hip_workspace_buffer(gs_hip_workspace* ws, int which, void** ptr, size_t* bytes) {
  return gs_hip_workspace_view_buffer(ws, 0, which, ptr, bytes);
}

int gs_hip_workspace_reserve_pairs(gs_hip_workspace* ws, uint64_t pairs) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "reserve_pairs: null workspace");
  return gsh::reserve_pairs(ws, pairs);
}

int gs_hip_workspace_load(gs_hip_workspace* ws, const float* params, const float* sh_rest, uint32_t count,
                          uint32_t sh_coeffs_in) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_load: null workspace");
  if (count > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_load: more slots than capacity");
  const size_t cap = ws->capacity, shn = cap * ws->sh_coeffs * 3;
  GS_HIP_TRY(hipDeviceSynchronize());
  int rc = 0;
  for (void* p : {(void*)ws->params, (void*)ws->m_params, (void*)ws->v_params}) {
    if (rc == 0) rc = gsh::fill_u32(ws, p, cap * 16, 0);
  }
  for (void* p : {(void*)ws->sh, (void*)ws->m_sh, (void*)ws->v_sh}) {
    if (rc == 0) rc = gsh::fill_u32(ws, p, shn, 0);
  }
  if (rc == 0) rc = gsh::fill_u32(ws, ws->stats, cap * 4, 0);
  if (rc != 0) return rc;
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  if (count == 0) return 0;
  GS_HIP_TRY(hipMemcpy(ws->params, params, size_t(count) * 64, hipMemcpyHostToDevice));
  if (ws->sh_coeffs > 0 && sh_rest != nullptr) {
    // [slot][coeff][rgb] -> planes [coeff][capacity][rgb]
    std::vector<float> planes(size_t(ws->sh_coeffs) * count * 3, 0.0f);
    const uint32_t used = std::min(sh_coeffs_in, ws->sh_coeffs);
    for (uint32_t i = 0; i < count; ++i) {
      for (uint32_t k = 0; k < used; ++k) {
        for (int c = 0; c < 3; ++c) {
          planes[(size_t(k) * count + i) * 3 + c] = sh_rest[(size_t(i) * sh_coeffs_in + k) * 3 + c];
        }
      }
    }
    for (uint32_t k = 0; k < ws->sh_coeffs; ++k) {
      GS_HIP_TRY(hipMemcpy(ws->sh + size_t(k) * cap * 3, planes.data() + size_t(k) * count * 3, size_t(count) * 12,
                           hipMemcpyHostToDevice));
    }
  }
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_workspace_store(gs_hip_workspace* ws, float* params, float* sh_rest, uint32_t count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_store: null workspace");
  if (count > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_store: more slots than capacity");
  GS_HIP_TRY(hipDeviceSynchronize());
  if (count == 0) return 0;
  if (params != nullptr) GS_HIP_TRY(hipMemcpy(params, ws->params, size_t(count) * 64, hipMemcpyDeviceToHost));
  if (sh_rest != nullptr && ws->sh_coeffs > 0) {
    std::vector<float> plane(size_t(count) * 3);
    for (uint32_t k = 0; k < ws->sh_coeffs; ++k) {
      GS_HIP_TRY(hipMemcpy(plane.data(), ws->sh + size_t(k) * ws->capacity * 3, plane.size() * 4,
                           hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < count; ++i) {
        for (int c = 0; c < 3; ++c) sh_rest[(size_t(i) * ws->sh_coeffs + k) * 3 + c] = plane[size_t(i) * 3 + c];
      }
    }
  }
  return 0;
}

int gs_hip_set_target(gs_hip_workspace* ws, const uint8_t* rgba8) {
  if (ws == nullptr || rgba8 == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_target: null");
  gsh::View& v = gsh::view0(ws);
  GS_HIP_TRY(hipMemcpyAsync(v.target, rgba8, size_t(ws->width) * ws->height * 4, hipMemcpyHostToDevice,
                            ws->stream));
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  return 0;
}

int gs_hip_set_target_device(gs_hip_workspace* ws, const void* device_rgba8) {
  if (ws == nullptr || device_rgba8 == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_target_device: null");
  gsh::View& v = gsh::view0(ws);
  GS_HIP_TRY(hipMemcpyAsync(v.target, device_rgba8, size_t(ws->width) * ws->height * 4, hipMemcpyDeviceToDevice,
                            ws->stream));
  return 0;
}

int gs_hip_workspace_get_tuning(gs_hip_workspace* ws, gs_hip_tuning* out) {
  if (ws == nullptr || out == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "get_tuning: null");
  *out = ws->tuning;
  return 0;
}

int gs_hip_workspace_set_tuning(gs_hip_workspace* ws, const gs_hip_tuning* t) {
  if (ws == nullptr || t == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_tuning: null");
  if (t->depth_digit_bits < 4 || t->depth_digit_bits > 8 ||
      (t->tile_digit_bits != 0 && (t->tile_digit_bits < 4 || t->tile_digit_bits > 8))) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "set_tuning: digit widths are 4..8 bits (tiles: 0 for one counting pass)");
  }
  if (t->raster_groups_per_cu > 64) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_tuning: at most 64 groups per CU");
  if (t->bin_shift > 2) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_tuning: bin_shift is 0..2");
  if (t->stored_state > 1) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_tuning: stored_state is 0 or 1");
  const bool changed = std::memcmp(&ws->tuning, t, sizeof(*t)) != 0;
  ws->tuning = *t;
  if (changed) gsh::drop_graph(ws);
  return 0;
}

int gs_hip_workspace_overflows(gs_hip_workspace* ws, uint64_t* steps) {
  if (ws == nullptr || steps == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "overflows: null");
  *steps = ws->overflow_steps;
  return 0;
}

int gs_hip_workspace_timings(gs_hip_workspace* ws, int enable, const char** names, double* ms, int capacity,
                             int* count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "timings: null workspace");
  gsh::Timer& t = ws->timer;
  int n = 0;
  if (t.enabled && t.used > 0) {
    GS_HIP_TRY(hipDeviceSynchronize());
    t.ms.assign(t.used, 0.0);
    for (size_t i = 0; i < t.used; ++i) {
      float v = 0.0f;
      if (hipEventElapsedTime(&v, t.start[i], t.stop[i]) == hipSuccess) t.ms[i] = v;
    }
    n = int(std::min<size_t>(t.used, size_t(std::max(capacity, 0))));
    for (int i = 0; i < n; ++i) {
      if (names != nullptr) names[i] = t.names[i].c_str();
      if (ms != nullptr) ms[i] = t.ms[i];
    }
  }
  if (count != nullptr) *count = n;
  t.used = 0;
  t.enabled = enable != 0;
  return 0;
}

}  // extern "C"
