// Host-side internals shared by the C ABI translation units: the workspace
// (the model's per-slot state plus one or more views' transient buffers),
// error reporting, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gs_hip.h"

namespace gsh {

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

// Radix sort scratch for up to `items` (key, value) pairs.
struct SortScratch {
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint32_t* hist = nullptr;          // [digits][blocks]
  uint32_t* digit_totals = nullptr;  // [digits]
  uint64_t items = 0;
};

// One view's transient buffers and its stream: everything a view-step
// computes before the optimizer, so several views can be in flight at once.
struct View {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // per slot
  float* records = nullptr;
  float* footprint = nullptr;
  uint32_t* tiles = nullptr;
  uint32_t* misc = nullptr;  // per slot: radius << 3 | colour clamp bits (the adjoint's view of the record)
  uint32_t* depth_keys = nullptr;
  uint32_t* depth_order = nullptr;
  uint32_t* offsets = nullptr;
  float* grad2d = nullptr;
  float* sh_input = nullptr;     // per slot: 8 floats, the view's SH colour gradient and direction
  uint32_t* flags = nullptr;     // densify: per-slot actions
  uint32_t* scan_tmp = nullptr;  // per-slot scan output
  // per pair
  uint64_t pair_capacity = 0;
  uint32_t* pair_tiles = nullptr;
  uint32_t* pair_gaussians = nullptr;
  uint64_t pairs = 0;  // of the last bin
  // per tile / pixel
  uint32_t* ranges = nullptr;
  uint32_t* tile_order = nullptr;  // tiles, longest list first (launch order of the rasterizers)
  float* image = nullptr;
  uint32_t* last = nullptr;
  uint32_t* target = nullptr;
  float* dldc = nullptr;
  float* ssim_maps = nullptr;  // float[9][pixels]: A, B, C per channel
  float* tile_loss = nullptr;
  uint32_t* totals = nullptr;  // [0] pairs, [1] emit mismatches, [5..7] densify counters
  uint32_t* block_sums = nullptr;
  uint32_t block_sums_entries = 0;
  uint32_t* host_totals = nullptr;  // pinned
  float* host_tile_loss = nullptr;  // pinned
  SortScratch sort;
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
  // the model: per-slot parameters, Adam moments, densify statistics
  float* params = nullptr;
  float* sh = nullptr;
  float* m_params = nullptr;
  float* v_params = nullptr;
  float* m_sh = nullptr;
  float* v_sh = nullptr;
  float* stats = nullptr;
  std::vector<gsh::View*> views;  // views[0] always exists; its stream runs the optimizer
  gsh::Timer timer;
};

namespace gsh {
inline View& view0(gs_hip_workspace* ws) { return *ws->views[0]; }
// Kernel timing on a view's stream (no-ops unless enabled).
void timer_begin(gs_hip_workspace* ws, View& v, const char* name);
void timer_end(gs_hip_workspace* ws, View& v);
// Allocates / frees one view's buffers.
int create_view(gs_hip_workspace* ws, View** out);
void destroy_view(View* v);
int reserve_pairs(gs_hip_workspace* ws, View& v, uint64_t pairs);
// Grows v.sort for `items`.
int reserve_sort(View& v, uint64_t items);
// Stable LSD radix sort of (keys, vals)[0, n) by key bits [0, bits) on v.stream.
int radix_sort(View& v, uint32_t* keys, uint32_t* vals, uint32_t n, int bits);
// Exclusive scan of `count` u32 values (values[gather[k]] when gather is not
// null) into out; the total goes to *total_dev (device).
int exclusive_scan(View& v, const uint32_t* values, const uint32_t* gather, uint32_t* out, uint32_t count,
                   uint32_t* total_dev);
// Stages on one view.
int launch_project(gs_hip_workspace* ws, View& v, const gs_hip_frame* f);
int bin_count(gs_hip_workspace* ws, View& v, const gs_hip_frame* f);          // async; P lands in v.host_totals
int bin_pairs(gs_hip_workspace* ws, View& v, const gs_hip_frame* f, uint64_t pairs);
int launch_forward(gs_hip_workspace* ws, View& v, const gs_hip_frame* f);
int launch_loss(gs_hip_workspace* ws, View& v, const gs_hip_frame* f, bool copy_losses);
int launch_backward(gs_hip_workspace* ws, View& v, const gs_hip_frame* f);
// Adjoint + Adam over the mean gradient of views[0, count) on view 0's stream.
int launch_adam(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count);
// Sums a view's tile losses (after its stream has finished).
void sum_losses(gs_hip_workspace* ws, View& v, double losses[3]);
}  // namespace gsh
