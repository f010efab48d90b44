// Host-side internals shared by the C ABI translation units: the workspace
// (all device buffers of one trainer), error reporting, launch helpers.
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

#define GS_HIP_TRY(expr)                                   \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return ::gsh::hip_error(_e, #expr); \
  } while (0)

#define GS_CHECK_LAUNCH(name)                               \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return ::gsh::hip_error(_e, name); \
  } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

// Radix sort scratch for up to `items` (key, value) pairs.
struct SortScratch {
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint32_t* hist = nullptr;        // [digits][blocks]
  uint32_t* digit_totals = nullptr;  // [digits]
  uint64_t items = 0;
  uint32_t hist_entries = 0;
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
  hipStream_t stream = nullptr;

  // per slot
  float* params = nullptr;
  float* sh = nullptr;
  float* m_params = nullptr;
  float* v_params = nullptr;
  float* m_sh = nullptr;
  float* v_sh = nullptr;
  float* stats = nullptr;
  float* records = nullptr;
  float* footprint = nullptr;
  uint32_t* tiles = nullptr;
  uint32_t* depth_keys = nullptr;
  uint32_t* depth_order = nullptr;
  uint32_t* offsets = nullptr;
  float* grad2d = nullptr;
  uint32_t* flags = nullptr;      // densify: per-slot action / release flags
  uint32_t* scan_tmp = nullptr;   // per slot scan scratch
  // per pair
  uint64_t pair_capacity = 0;
  uint32_t* pair_tiles = nullptr;
  uint32_t* pair_gaussians = nullptr;
  // per tile / pixel
  uint32_t* ranges = nullptr;
  float* image = nullptr;
  uint32_t* last = nullptr;
  uint32_t* target = nullptr;
  float* dldc = nullptr;
  float* ssim_maps = nullptr;  // float[9][pixels]: A, B, C per channel
  float* tile_loss = nullptr;
  uint32_t* totals = nullptr;  // [0] pairs, [1..3] densify counters
  uint32_t* block_sums = nullptr;
  uint32_t block_sums_entries = 0;
  uint32_t* host_totals = nullptr;  // pinned
  float* host_tile_loss = nullptr;  // pinned
  gsh::SortScratch sort;
  gsh::Timer timer;
};

namespace gsh {
// Kernel timing (no-ops unless enabled).
void timer_begin(gs_hip_workspace* ws, const char* name);
void timer_end(gs_hip_workspace* ws);
// Grows ws->sort for `items`.
int reserve_sort(gs_hip_workspace* ws, uint64_t items);
// Stable LSD radix sort of (keys, vals)[0, n) by key bits [0, bits) on ws->stream;
// the result ends in keys / vals.
int radix_sort(gs_hip_workspace* ws, uint32_t* keys, uint32_t* vals, uint32_t n, int bits);
// Exclusive scan of `count` u32 values (values[gather[k]] when gather is not
// null) into out; the total goes to *total_dev (device).
int exclusive_scan(gs_hip_workspace* ws, const uint32_t* values, const uint32_t* gather, uint32_t* out,
                   uint32_t count, uint32_t* total_dev);
}  // namespace gsh
