// Host-side internals shared by the C ABI translation units: the workspace
// (the model's per-slot state plus up to GS_HIP_MAX_VIEWS views' transient
// buffers), the device tables the batched kernels read, error reporting and
// launch helpers.
//
// Every stage is one launch over all views of a step (the view is blockIdx.y,
// or part of the work item for the rasterizers), on one stream.  Kernels find
// their view's frame and buffers in a device-resident table (ViewDev) and
// their scene's model in another (SceneDev), so the step's launch sequence
// does not depend on the frames and can be captured once in a HIP graph and
// replayed.  A step's views may belong to one scene (a workspace) or to
// several (a batch, whose launch context borrows the scenes' views): grids
// are sized for the largest scene, pair buffer and tile grid of the step
// (StepShape) and each block finds its own extent in the tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gs_hip.h"

namespace gsh {

// The SH optimizer's split of the SH rest coefficients: one wave per group
// of kShGroup coefficients of 64 slots (its planes, colour gradients and
// Adam moments in registers), kShGroups groups at degree 3.
constexpr int kShGroup = 5;
constexpr int kShGroups = (GS_HIP_MAX_SH_REST + kShGroup - 1) / kShGroup;

// Thread-local last error; every C entry point returns its code.
int set_error(int code, const std::string& message);
int hip_error(hipError_t e, const char* what);
void clear_error();

#define GS_HIP_TRY(expr)                                      \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return ::gsh::hip_error(_e, #expr); \
  } while (0)

#define GS_CHECK_LAUNCH(name)                                \
  do {                                                       \
    hipError_t _e = hipGetLastError();                       \
    if (_e != hipSuccess) return ::gsh::hip_error(_e, name); \
  } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

// One segment of a batched radix sort: (keys, vals)[0, n) with
// n = min(*count, count_max) (count_max when count is null), sorted in place
// through the alternates.
struct SortSeg {
  uint32_t* keys;
  uint32_t* vals;
  uint32_t* keys_alt;
  uint32_t* vals_alt;
  uint32_t* hist;          // [256][blocks]
  uint32_t* digit_totals;  // [256]
  const uint32_t* count;
  uint32_t count_max;
  uint32_t blocks;  // 4096-item blocks the scratch holds (the hist row length); a launch's extra blocks idle
};

// One segment of a batched exclusive scan: out[i] = sum of values[gather[j]]
// (or values[j] >> shift) for j < i, i < n = min(*count, count_max)
// (count_max when count is null); the total goes to *total.
struct ScanSeg {
  const uint32_t* values;
  const uint32_t* gather;
  uint32_t* out;
  uint32_t* block_sums;
  uint32_t* total;
  const uint32_t* count;
  uint32_t count_max;
  uint32_t blocks;  // block_sums entries; a launch's extra blocks idle
  uint32_t shift;
};

// A scene's model as the per-slot kernels see it, and its views in the step's
// view table: views [first_view, first_view + views).
struct SceneDev {
  float4* params;
  float* sh;
  float4* m_params;
  float4* v_params;
  float* m_sh;
  float* v_sh;
  float4* stats;
  float4* dir_grad;  // [kShGroups][capacity]: per group of kShGroup SH rest coefficients, its view-direction part of dL/dposition
  uint32_t capacity;
  uint32_t sh_coeffs;
  uint32_t first_view;
  uint32_t views;
};

// A view's frame and buffers as the kernels see them (device memory).
struct ViewDev {
  gs_hip_frame f;
  // per slot
  float4* records;
  float4* footprint;  // [capacity][2]: (cov abc, threshold), (mean xy, tiles bits, 0)
  uint32_t* tiles;
  uint32_t* misc;  // radius << 3 | colour clamp bits: what the adjoint needs of the record
  uint32_t* depth_keys;
  uint32_t* depth_order;
  uint32_t* offsets;
  float* grad2d;  // [capacity][9]
  unsigned long long* grad_fx;  // [capacity][9] fixed-point sums of the deterministic backward, or null
  // per pair
  uint32_t* pair_tiles;
  uint32_t* pair_gaussians;
  // where emission writes the pairs: the pair buffers (the radix tile sort
  // sorts them in place) or the sort's alternates (the counting tile sort
  // scatters them from there into the pair buffers)
  uint32_t* emit_tiles;
  uint32_t* emit_gaussians;
  uint32_t* tile_hist;  // counting tile sort: [tile_cols][tile_count + 1] per-chunk tile counts
  uint32_t tile_cols;
  // per tile / pixel
  uint2* ranges;
  float4* image;
  uint32_t* last;
  const uint32_t* target;
  float4* dldc;
  float* maps;  // [9][pixels]: SSIM gradient maps A, B, C per channel
  float4* tile_loss;
  float4* tile_fit;   // [tiles][2]: exposure fit sums (GS_HIP_FLAG_EXPOSURE)
  float4* pose_grad;  // [capacity / 256][2]: pose gradient partials (GS_HIP_FLAG_POSE_GRAD)
  uint32_t* totals;  // [0] pairs, [1] emit mismatches, [5..7] densify counters, [8] visible splats (viewer)
  uint64_t pair_capacity;
  uint32_t capacity;    // the scene's slot capacity
  uint32_t tile_count;  // tiles_x * tiles_y (16x16 raster tiles)
  // Pair bins: squares of (16 << bin_shift) pixels; pair keys, ranges and
  // the TILES counts are per bin, and raster tile (tx, ty) walks the list of
  // bin (tx >> bin_shift, ty >> bin_shift).
  uint32_t bin_shift;
  uint32_t bins_x;
  uint32_t bin_count;  // bins_x * bins_y
  // The projection's ray clamp, 1.3 max(cx, W - cx) / fx and likewise in y,
  // in IEEE single precision like the device expression it replaces.
  float lim_x, lim_y;
  // Stored-state backward: per-pixel (C rgb, T) at the start of every 64-pair
  // bucket after a tile's first, kBucketPixels float4 per bucket slot (null
  // when the step re-walks); see bucket_slot.
  float4* bucket_state;
  uint64_t bucket_slots;
  // The launch context's flag (k_work_count's): some view of the step found
  // more pairs than its buffers hold.
  const uint32_t* step_overflow;
};

// The launch context's work counters: [0..1] raster work, [2..3] stored-state
// buckets, [4] the step's pair-overflow flag (set by k_work_count once the
// pair counts are known), [5] the work order's finished blocks, then its
// count and taken places of each of its kWorkBuckets list-length buckets.
constexpr uint32_t kStepOverflowCounter = 4;
constexpr uint32_t kWorkDoneCounter = 5;
constexpr uint32_t kWorkBuckets = 33;
constexpr uint32_t kWorkBucketCounters = 16;
constexpr uint32_t kWorkCounters = kWorkBucketCounters + 2 * kWorkBuckets;

// Stored-state backward: list positions per bucket (one wave's splats) and
// the pixels of a bucket slot (a 16x16 tile, quadrant-major: pixel 64 q + l
// is lane l of the wave that owns quadrant q).
constexpr uint32_t kBucketSplats = 64;
constexpr uint32_t kBucketPixels = 256;

// The slot of bucket k >= 1 of raster tile `tile` walking the list of bin
// `bin` that starts at pair `range_x`.  Bins' lists partition the pairs, so
// with L = the bin's length, (range_x / 64 + bin) grows by at least
// floor(L / 64) + 1 from one bin to the next, more than the L / 64 buckets
// after the first that its list can need: a bin's slots never meet the
// next's, and every tile of the bin gets its own column of the 4^bin_shift.
__host__ __device__ inline uint64_t bucket_slot(uint32_t range_x, uint32_t bin, uint32_t k, uint32_t tile,
                                                uint32_t tiles_x, uint32_t bin_shift) {
  const uint32_t mask = (1u << bin_shift) - 1u;
  const uint32_t sub = ((tile / tiles_x) & mask) << bin_shift | ((tile % tiles_x) & mask);
  return ((uint64_t(range_x / kBucketSplats) + bin + k - 1u) << (2u * bin_shift)) + sub;
}
// Slots a view needs for up to `pairs` pairs in `bins` bins.
__host__ __device__ inline uint64_t bucket_slots_for(uint64_t pairs, uint32_t bins, uint32_t bin_shift) {
  return (pairs / kBucketSplats + bins + 1u) << (2u * bin_shift);
}
// A bucket of the stored-state backward: which (view, tile) and which 64 pairs.
struct BucketEntry {
  uint32_t item;  // view << kTileBits | tile
  uint32_t k;     // list positions [64 k, 64 k + 64) of the tile's walk
};

// The bin raster tile `tile` walks.
__host__ __device__ inline uint32_t bin_of_tile(const ViewDev& v, uint32_t tile) {
  const uint32_t tx = tile % v.f.tiles_x, ty = tile / v.f.tiles_x;
  return (ty >> v.bin_shift) * v.bins_x + (tx >> v.bin_shift);
}

// Rasterizer work items name (view, tile): tiles per view stay below 2^16.
constexpr uint32_t kTileBits = 16;
constexpr uint32_t kMaxTilesPerView = (1u << kTileBits) - 1u;
// A rasterizer work item for a quad member outside the tile grid.
constexpr uint32_t kNoWork = 0xffffffffu;

// Rasterizer work items of a view: its tiles in 2x2 quads (the quads of the
// last column and row may stick out of the grid).
__host__ __device__ inline uint32_t quads_x(uint32_t tiles_x) { return (tiles_x + 1u) / 2u; }
__host__ __device__ inline uint32_t quad_items(uint32_t tiles_x, uint32_t tiles_y) {
  return 4u * quads_x(tiles_x) * ((tiles_y + 1u) / 2u);
}
// A launch context's table sizes: a workspace's own views, or a batch's.
constexpr uint32_t kMaxBatchViews = 64;
constexpr uint32_t kMaxBatchScenes = 32;

// The extent of the current step's table, set when it is uploaded.
struct StepShape {
  uint32_t views = 0;
  uint32_t scenes = 0;
  uint32_t max_capacity = 0;       // the largest scene's slot capacity (per-slot grids)
  uint64_t max_pair_capacity = 0;  // the largest pair buffer (per-pair grids)
  uint32_t max_tiles = 0;          // the largest tile grid (per-tile grids)
  uint32_t total_tiles = 0;        // every view's tiles (per-tile loss entries)
  uint32_t work_items = 0;         // rasterizer work items: 4 per 2x2 quad of tiles of every view
  uint32_t max_quads = 0;          // the most 2x2 quads of a view
  uint32_t max_bins = 0;           // the most pair bins of a view
  uint32_t tile_key_bits = 0;      // tile sort key width (bins)
  bool counting_tile_sort = false; // one counting pass (tuning.tile_digit_bits == 0 and the tile grids fit)
  bool any_sh = false;             // some scene stores SH rest coefficients
  bool any_pose = false;           // some view asks for its pose gradient
  bool deterministic = false;      // some view asks for GS_HIP_FLAG_DETERMINISTIC (then every view's backward is)
  bool stored_state = false;       // the stored-state backward runs (tuning.stored_state, not deterministic)
};

// Radix sort scratch of one view, for up to `items` (key, value) pairs.
struct SortScratch {
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint32_t* hist = nullptr;          // [256][blocks]
  uint32_t* digit_totals = nullptr;  // [256]
  uint64_t items = 0;
  uint32_t* tile_hist = nullptr;     // counting tile sort: [tile_cols][tiles + 1]
  uint32_t tile_cols = 0;            // chunks of kTileChunk pairs the table holds
};

// Counting tile sort: pairs per chunk (one wave ranks a chunk), and the
// largest tile grid (+1 for the bin of a slot whose tiles were miscounted)
// its per-chunk LDS counters hold.
constexpr uint32_t kTileChunk = 16384;
constexpr uint32_t kCountingMaxBins = 16384;

// One view's transient buffers (host handles of what ViewDev points at).
struct View {
  float* records = nullptr;
  float* footprint = nullptr;
  uint32_t* tiles = nullptr;
  uint32_t* misc = nullptr;
  uint32_t* depth_keys = nullptr;
  uint32_t* depth_order = nullptr;
  uint32_t* offsets = nullptr;
  float* grad2d = nullptr;
  unsigned long long* grad_fx = nullptr;  // deterministic backward: [capacity][9], made on first use
  float* bucket_state = nullptr;  // stored-state backward: [bucket_slots][256][4], made on first use
  uint64_t bucket_slots = 0;
  uint32_t* flags = nullptr;     // densify: per-slot actions
  uint32_t* scan_tmp = nullptr;  // per-slot scan output
  uint64_t pair_capacity = 0;
  uint32_t* pair_tiles = nullptr;
  uint32_t* pair_gaussians = nullptr;
  uint64_t pairs = 0;  // of the last step (host copy)
  uint32_t* ranges = nullptr;
  float* image = nullptr;
  uint32_t* last = nullptr;
  uint32_t* target = nullptr;
  float* dldc = nullptr;
  float* ssim_maps = nullptr;
  float* tile_loss = nullptr;  // slice of the workspace's tile_loss_all
  float* tile_fit = nullptr;   // [tiles][8]
  float* pose_grad = nullptr;  // [div_up(capacity, 256)][8]
  uint32_t* totals = nullptr;  // slice of the workspace's totals_all
  uint32_t* block_sums = nullptr;
  uint32_t block_sums_entries = 0;
  SortScratch sort;
  uint32_t capacity = 0;  // of the scene the view belongs to
  uint32_t tile_grid = 0;  // tiles_x * tiles_y of the view's image
};

struct Timer {
  bool enabled = false;
  std::vector<std::string> names;
  std::vector<hipEvent_t> start, stop;
  size_t used = 0;
  std::vector<double> ms;
};

}  // namespace gsh

struct gs_hip_workspace {
  uint32_t capacity = 0;
  int sh_degree = 0;
  uint32_t sh_coeffs = 0;  // rest coefficients stored per slot
  uint32_t width = 0, height = 0, tiles_x = 0, tiles_y = 0;
  uint32_t cus = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  // the model: per-slot parameters, Adam moments, densify statistics
  float* params = nullptr;
  float* sh = nullptr;
  float* m_params = nullptr;
  float* v_params = nullptr;
  float* m_sh = nullptr;
  float* v_sh = nullptr;
  float* stats = nullptr;
  float* dir_grad = nullptr;  // [kShGroups][capacity][4] (SH rest stored)
  std::vector<gsh::View*> views;  // views[0] always exists (a batch's launch context borrows its scenes' views)
  bool borrowed_views = false;
  // device tables: ViewDev[table_views], SceneDev[table_scenes] and the batched sort / scan segments
  uint32_t table_views = GS_HIP_MAX_VIEWS, table_scenes = 1;
  gsh::ViewDev* dev_views = nullptr;
  gsh::SceneDev* dev_scenes = nullptr;
  gsh::SortSeg* dev_depth_sort = nullptr;  // [table_views]
  gsh::SortSeg* dev_tile_sort = nullptr;   // [table_views]
  gsh::ScanSeg* dev_tile_scan = nullptr;   // [table_views]
  gsh::ScanSeg* dev_scan_tmp = nullptr;    // [1]: host-count scans (densify)
  uint32_t* work = nullptr;           // rasterizer work items (view << 16 | tile), longest lists first
  uint32_t work_entries = 0;
  uint32_t* work_counters = nullptr;  // kWorkCounters: [0] forward, [1] backward (persistent rasterizers), [2] buckets
                                      // listed, [3] taken, [4] the step's pair overflow, [16..] the work order's buckets
  gsh::BucketEntry* bucket_list = nullptr;  // stored-state backward: the step's buckets, listed by the forward
  uint64_t bucket_list_entries = 0;
  uint32_t* totals_all = nullptr;     // [table_views][16]
  float* tile_loss_all = nullptr;     // [tile_loss_entries][4], view v's tiles from tile_base[v]
  uint32_t tile_loss_entries = 0;
  std::vector<uint32_t> tile_base;    // per view of the current table
  gsh::StepShape shape;               // of the current table
  // host side of the tables and readbacks (pinned)
  gsh::ViewDev* host_views = nullptr;    // staging ring: kStagingSlots copies of ViewDev[table_views]
  gsh::SceneDev* host_scenes = nullptr;  // and of SceneDev[table_scenes]
  std::vector<hipEvent_t> staging_done;
  uint32_t staging_next = 0;
  uint32_t* host_totals = nullptr;  // [table_views][16]
  float* host_tile_loss = nullptr;  // [tile_loss_entries][4]
  uint64_t table_signature = 0;     // buffers the segments were built for (0: rebuild)
  uint64_t max_pairs_seen = 0;
  uint64_t overflow_steps = 0;
  gs_hip_tuning tuning;
  // captured batched step
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_signature = 0;  // table, view count and step kind it was captured for
  gsh::Timer timer;
};

namespace gsh {
constexpr uint32_t kStagingSlots = 64;

inline View& view0(gs_hip_workspace* ws) { return *ws->views[0]; }
inline uint32_t tile_count(const gs_hip_workspace* ws) { return ws->tiles_x * ws->tiles_y; }

// One view of a step's table: whose buffers, which frame, which scene's model.
struct TableEntry {
  View* view;
  const gs_hip_frame* frame;
  gs_hip_workspace* scene;
  const void* target;  // device RGBA8 read in place, or null for the view's TARGET buffer
};
// Kernel timing on the workspace stream (no-ops unless enabled).
void timer_begin(gs_hip_workspace* ws, const char* name);
void timer_end(gs_hip_workspace* ws);
// Destroys the timing events (at the workspace's or batch's end).
void release_timer(Timer& t);
// Allocates / frees one view's buffers.
int create_view(gs_hip_workspace* ws, uint32_t index, View** out);
void destroy_view(View* v);
// Grows every view's pair buffers (and sort scratch) to hold `pairs` with
// headroom; drops the captured graph when buffers move.
int reserve_pairs(gs_hip_workspace* ws, uint64_t pairs);
int reserve_sort(View& v, uint64_t items);
// Writes frames[0, count) into the device table for views [0, count)
// (stream-ordered copy from a pinned staging slot) with every view's buffer
// pointers, and the sort / scan segments when the buffers moved.
// targets (may be NULL, as may its entries): device RGBA8 photos the loss
// kernels read in place instead of the views' own TARGET buffers.
int upload_table(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count,
                 const void* const* targets = nullptr, int bin_shift = -1);
// The same for any entries (a batch's views of several scenes, grouped by
// scene): sets ws->shape and ws->tile_base.
// bin_shift < 0: the launch context's tuning.bin_shift.
int upload_entries(gs_hip_workspace* ws, const std::vector<TableEntry>& entries, int bin_shift = -1);
// Batched stages over views [0, count) of the uploaded table.
int launch_project(gs_hip_workspace* ws, uint32_t count);
int launch_depth_sort_and_scan(gs_hip_workspace* ws, uint32_t count, bool depth16);
int launch_emit_and_tile_sort(gs_hip_workspace* ws, uint32_t count);
// Pieces of the above for other emitters: ranges cleared (and the mismatch
// counter), then the stable tile sort, ranges and the rasterizers' work order.
int launch_clear_ranges(gs_hip_workspace* ws, uint32_t count);
int launch_tile_sort_and_ranges(gs_hip_workspace* ws, uint32_t count);
int launch_forward(gs_hip_workspace* ws, uint32_t count);
int launch_loss(gs_hip_workspace* ws, uint32_t count);
int launch_backward(gs_hip_workspace* ws, uint32_t count);
// Adjoint + Adam over the mean gradient of views [0, count); skipped on the
// device when any view's pairs overflowed its buffers.
int launch_adam(gs_hip_workspace* ws, uint32_t count);
// Every stage of one step over the uploaded table, eagerly or by replaying
// the captured graph (tuning.graph), ending with the totals' readback.
int launch_step(gs_hip_workspace* ws, uint32_t count, bool depth16, bool train, bool losses);
// Device tables of a launch context for `views` views of up to `scenes`
// scenes and `tile_entries` tiles in all; and their release.
int alloc_tables(gs_hip_workspace* ws, uint32_t views, uint32_t scenes, uint32_t tile_entries);
void release_tables(gs_hip_workspace* ws);
// One batched step over views [0, count) (train: with backward and Adam;
// losses: tile loss sums read back), rerun with larger pair buffers when a
// view overflowed; pairs[count] and loss_sums[3 * count] may be null.
int run_step(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets, bool train,
             bool losses, uint64_t* pairs, double* loss_sums);
// Copies the views' totals (and tile losses) to the pinned host copies.
int launch_readback(gs_hip_workspace* ws, uint32_t count, bool losses);
// Exclusive scan of `count` u32 on the stream (one segment, host count).
int exclusive_scan(gs_hip_workspace* ws, View& v, const uint32_t* values, const uint32_t* gather, uint32_t* out,
                   uint32_t count, uint32_t* total_dev);
// Sums view v's tile losses (host copy, after the stream finished).
void sum_losses(gs_hip_workspace* ws, uint32_t v, double losses[3]);
// Fills `words` u32 with `value` by a kernel (graph-safe, unlike hipMemsetAsync).
int fill_u32(gs_hip_workspace* ws, void* p, uint64_t words, uint32_t value);
// Drops the captured graph (buffers or view count changed).
void drop_graph(gs_hip_workspace* ws);
}  // namespace gsh
