// Binning: depth sort of the slots, exclusive scan of their tile counts in
// depth order, pair emission, stable tile sort of the pairs and per-tile
// ranges.  The order this produces is that of sorting 64-bit keys
// (tile << 32) | depth_key with ties broken by the gaussian index: the slots
// are sorted stably by depth key, each emits its pairs at its offset in that
// order, and a stable sort by tile alone then leaves each tile's pairs front
// to back (engine/src/trainer.cpp:665-711, train_emit.comp:3-9).
//
// The radix sort is LSD with up to 8-bit digits.  Per pass: an upsweep kernel
// counts digits per block of 4096 items (per-wave LDS histograms), a scan
// kernel turns the [digit][block] counts into offsets, and a scatter kernel
// ranks every item stably inside its block with wave64 ballots (the lanes
// holding the same digit, and the rank as the popcount of those below) plus
// per-wave digit counts in LDS.
#include <algorithm>

#include "gs_common.h"
#include "gs_internal.h"

namespace {

using namespace gsd;

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortItems = kSortThreads * kSortRounds;  // 4096 per block
constexpr int kWaves = kSortThreads / 64;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Lanes of this wave whose `digit` (nbits wide) equals this lane's, among `active`.
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, int nbits, uint64_t active) {
  uint64_t peers = active;
  for (int b = 0; b < nbits; ++b) {
    const bool bit = ((digit >> b) & 1u) != 0u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__global__ void __launch_bounds__(kSortThreads) k_radix_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                                int shift, int nbits, uint32_t* __restrict__ hist,
                                                                uint32_t blocks) {
  __shared__ uint32_t h[kWaves][256];
  const uint32_t radix = 1u << nbits, mask = radix - 1u;
  const int wave = threadIdx.x / 64;
  for (int d = threadIdx.x; d < kWaves * 256; d += kSortThreads) (&h[0][0])[d] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortItems;
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t idx = base + r * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[wave][(keys[idx] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) {
    uint32_t s = 0;
    for (int w = 0; w < kWaves; ++w) s += h[w][d];
    hist[size_t(d) * blocks + blockIdx.x] = s;
  }
}

// One block per digit: exclusive scan of that digit's per-block counts, and
// the digit's total.
__global__ void __launch_bounds__(256) k_radix_scan(uint32_t* __restrict__ hist, uint32_t blocks,
                                                    uint32_t* __restrict__ digit_totals) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_carry;
  uint32_t* row = hist + size_t(blockIdx.x) * blocks;
  if (threadIdx.x == 0) s_carry = 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t start = 0; start < blocks; start += 256) {
    const uint32_t idx = start + threadIdx.x;
    const uint32_t v = idx < blocks ? row[idx] : 0u;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    if (idx < blocks) row[idx] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) digit_totals[blockIdx.x] = s_carry;
}

__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, uint32_t n, int shift,
                                                                int nbits, const uint32_t* __restrict__ hist,
                                                                const uint32_t* __restrict__ digit_totals,
                                                                uint32_t blocks) {
  __shared__ uint32_t s_running[256];
  __shared__ uint32_t s_counts[kWaves][256];
  const uint32_t radix = 1u << nbits, mask = radix - 1u;
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  // Start of each digit: the digits below it (a block scan of the totals),
  // plus this digit's items in the blocks before this one.
  {
    const uint32_t v = threadIdx.x < radix ? digit_totals[threadIdx.x] : 0u;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_counts[0][wave] = incl;
    __syncthreads();
    uint32_t below = incl - v;
    for (int w = 0; w < wave; ++w) below += s_counts[0][w];
    s_running[threadIdx.x] = threadIdx.x < radix ? below + hist[size_t(threadIdx.x) * blocks + blockIdx.x] : 0u;
    __syncthreads();
  }
  for (int d = threadIdx.x; d < kWaves * 256; d += kSortThreads) (&s_counts[0][0])[d] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortItems;
  const uint64_t below_me = lanemask_lt();
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t idx = base + r * kSortThreads + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t key = valid ? keys_in[idx] : 0u;
    const uint32_t val = valid ? vals_in[idx] : 0u;
    const uint32_t digit = (key >> shift) & mask;
    const uint64_t peers = match_digit(digit, nbits, __ballot(valid));
    const uint32_t rank = __popcll(peers & below_me);
    if (valid && rank == 0u) s_counts[wave][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_running[digit] + rank;
      for (int w = 0; w < wave; ++w) pos += s_counts[w][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    if (threadIdx.x < radix) {
      uint32_t s = 0;
      for (int w = 0; w < kWaves; ++w) {
        s += s_counts[w][threadIdx.x];
        s_counts[w][threadIdx.x] = 0u;
      }
      s_running[threadIdx.x] += s;
    }
    __syncthreads();
  }
}

// ---- exclusive scan ----------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanRounds = 16;
constexpr int kScanItems = kScanThreads * kScanRounds;

__device__ __forceinline__ uint32_t load_value(const uint32_t* __restrict__ values, const uint32_t* __restrict__ gather,
                                               uint32_t idx) {
  return gather != nullptr ? values[gather[idx]] : values[idx];
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* __restrict__ values,
                                                              const uint32_t* __restrict__ gather, uint32_t count,
                                                              uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s[kScanThreads / 64];
  uint32_t sum = 0;
  const uint32_t base = blockIdx.x * kScanItems;
  for (int r = 0; r < kScanRounds; ++r) {
    const uint32_t idx = base + r * kScanThreads + threadIdx.x;
    if (idx < count) sum += load_value(values, gather, idx);
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) block_sums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t* __restrict__ block_sums, uint32_t blocks,
                                                      uint32_t* __restrict__ total) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t start = 0; start < blocks; start += 1024) {
    const uint32_t idx = start + threadIdx.x;
    const uint32_t v = idx < blocks ? block_sums[idx] : 0u;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    if (idx < blocks) block_sums[idx] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int w = 0; w < 16; ++w) s += s_wave[w];
      s_carry += s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(const uint32_t* __restrict__ values,
                                                            const uint32_t* __restrict__ gather, uint32_t count,
                                                            const uint32_t* __restrict__ block_sums,
                                                            uint32_t* __restrict__ out) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = block_sums[blockIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kScanItems;
  for (int r = 0; r < kScanRounds; ++r) {
    const uint32_t idx = base + r * kScanThreads + threadIdx.x;
    const uint32_t v = idx < count ? load_value(values, gather, idx) : 0u;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    if (idx < count) out[idx] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
}

// ---- pair emission and ranges ------------------------------------------------
// One thread per slot in depth order: writes the slot's pairs (tile, slot) at
// its offset, row by row over the same footprint the projection counted
// (train_emit.comp:26-64).  A count that differs (it cannot: same code, same
// inputs) is padded with the unused tile `tiles` and counted in totals[1].
__global__ void __launch_bounds__(256) k_emit(gs_hip_frame f, const float4* __restrict__ records,
                                              const float4* __restrict__ footprint,
                                              const uint32_t* __restrict__ tiles,
                                              const uint32_t* __restrict__ order,
                                              const uint32_t* __restrict__ offsets, uint32_t* __restrict__ pair_tiles,
                                              uint32_t* __restrict__ pair_gaussians, uint64_t pair_capacity,
                                              uint32_t* __restrict__ totals) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= f.slots) return;
  const uint32_t i = order[k];
  const uint32_t count = tiles[i];
  if (count == 0u) return;
  const float4 r0 = records[size_t(i) * 3u];
  const float4 fp = footprint[i];
  uint64_t p = offsets[k];
  const uint64_t end = std::min<uint64_t>(p + count, pair_capacity);
  uint32_t wanted = 0;
  int row_first, row_last;
  footprint_rows(r0.y, fp.z, fp.w, f.cam.height, row_first, row_last);
  if (row_first <= row_last) {
    for (uint32_t ty = uint32_t(row_first) / kTile; ty <= uint32_t(row_last) / kTile; ++ty) {
      uint32_t x0, x1;
      footprint_row_tiles(r0.x, r0.y, fp.x, fp.y, fp.z, fp.w, row_first, row_last, ty, f.cam.width, x0, x1);
      wanted += x1 - x0;
      for (uint32_t x = x0; x < x1 && p < end; ++x, ++p) {
        pair_tiles[p] = ty * f.tiles_x + x;
        pair_gaussians[p] = i;
      }
    }
  }
  if (wanted == count) return;
  atomicAdd(&totals[1], 1u);
  for (; p < end; ++p) {
    pair_tiles[p] = f.tiles_x * f.tiles_y;
    pair_gaussians[p] = i;
  }
}

__global__ void __launch_bounds__(256) k_fill_ranges(uint32_t* __restrict__ ranges, uint32_t entries) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < entries) ranges[i] = 0u;
}

__global__ void __launch_bounds__(256) k_ranges(const uint32_t* __restrict__ keys, uint32_t count,
                                                uint2* __restrict__ ranges) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  const uint32_t tile = keys[s];
  if (s == 0u) {
    ranges[tile].x = 0u;
  } else {
    const uint32_t prev = keys[s - 1u];
    if (prev != tile) {
      ranges[prev].y = s;
      ranges[tile].x = s;
    }
  }
  if (s == count - 1u) ranges[tile].y = count;
}

// Launch order of the rasterizers: tiles bucketed by floor(log2(list
// length)), longest first, empty tiles last (longest-processing-time-first
// scheduling within a factor of two).  One workgroup.
__global__ void __launch_bounds__(1024) k_tile_order(const uint2* __restrict__ ranges, uint32_t tiles,
                                                     uint32_t* __restrict__ order) {
  __shared__ uint32_t s_count[33];
  if (threadIdx.x < 33) s_count[threadIdx.x] = 0u;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < tiles; t += blockDim.x) {
    const uint2 r = ranges[t];
    atomicAdd(&s_count[__clz(r.y - r.x)], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (int b = 0; b < 33; ++b) {
      const uint32_t c = s_count[b];
      s_count[b] = sum;
      sum += c;
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < tiles; t += blockDim.x) {
    const uint2 r = ranges[t];
    order[atomicAdd(&s_count[__clz(r.y - r.x)], 1u)] = t;
  }
}

int bits_for(uint32_t n) {
  int bits = 0;
  while (bits < 32 && (uint64_t(1) << bits) < n) ++bits;
  return bits;
}

}  // namespace

namespace gsh {

int reserve_sort(View& v, uint64_t items) {
  SortScratch& s = v.sort;
  if (items <= s.items) return 0;
  if (s.keys_alt) hipFree(s.keys_alt);
  if (s.vals_alt) hipFree(s.vals_alt);
  if (s.hist) hipFree(s.hist);
  if (s.digit_totals) hipFree(s.digit_totals);
  s = SortScratch{};
  const uint64_t blocks = (items + kSortItems - 1) / kSortItems;
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&s.keys_alt), items * 4));
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&s.vals_alt), items * 4));
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&s.hist), std::max<uint64_t>(blocks, 1) * 256 * 4));
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&s.digit_totals), 256 * 4));
  s.items = items;
  return 0;
}

int radix_sort(View& v, uint32_t* keys, uint32_t* vals, uint32_t n, int bits) {
  if (n <= 1 || bits <= 0) return 0;
  if (n > v.sort.items) return set_error(GS_HIP_E_CAPACITY, "radix_sort: scratch too small");
  const uint32_t blocks = div_up(n, kSortItems);
  const int passes = (bits + 7) / 8;
  uint32_t *ka = keys, *va = vals, *kb = v.sort.keys_alt, *vb = v.sort.vals_alt;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    const int nbits = std::min(8, bits - shift);
    k_radix_upsweep<<<blocks, kSortThreads, 0, v.stream>>>(ka, n, shift, nbits, v.sort.hist, blocks);
    GS_CHECK_LAUNCH("k_radix_upsweep");
    k_radix_scan<<<1u << nbits, 256, 0, v.stream>>>(v.sort.hist, blocks, v.sort.digit_totals);
    GS_CHECK_LAUNCH("k_radix_scan");
    k_radix_scatter<<<blocks, kSortThreads, 0, v.stream>>>(ka, va, kb, vb, n, shift, nbits, v.sort.hist,
                                                           v.sort.digit_totals, blocks);
    GS_CHECK_LAUNCH("k_radix_scatter");
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    GS_HIP_TRY(hipMemcpyAsync(keys, ka, size_t(n) * 4, hipMemcpyDeviceToDevice, v.stream));
    GS_HIP_TRY(hipMemcpyAsync(vals, va, size_t(n) * 4, hipMemcpyDeviceToDevice, v.stream));
  }
  return 0;
}

int exclusive_scan(View& v, const uint32_t* values, const uint32_t* gather, uint32_t* out, uint32_t count,
                   uint32_t* total_dev) {
  const uint32_t blocks = std::max<uint32_t>(div_up(count, kScanItems), 1);
  if (blocks > v.block_sums_entries) return set_error(GS_HIP_E_CAPACITY, "exclusive_scan: too many blocks");
  k_scan_reduce<<<blocks, kScanThreads, 0, v.stream>>>(values, gather, count, v.block_sums);
  GS_CHECK_LAUNCH("k_scan_reduce");
  k_scan_blocks<<<1, 1024, 0, v.stream>>>(v.block_sums, blocks, total_dev);
  GS_CHECK_LAUNCH("k_scan_blocks");
  k_scan_down<<<blocks, kScanThreads, 0, v.stream>>>(values, gather, count, v.block_sums, out);
  GS_CHECK_LAUNCH("k_scan_down");
  return 0;
}

// Depth sort and offsets; the pair count is copied to v.host_totals[0]
// asynchronously (valid once v.stream has reached this point).
int bin_count(gs_hip_workspace* ws, View& v, const gs_hip_frame* f) {
  if (f->slots == 0) {
    v.host_totals[0] = 0;
    return 0;
  }
  const int depth_bits = (f->flags & GS_HIP_FLAG_DEPTH16) != 0u ? 16 : 32;
  timer_begin(ws, v, "depth_sort");
  int rc = radix_sort(v, v.depth_keys, v.depth_order, f->slots, depth_bits);
  timer_end(ws, v);
  if (rc != 0) return rc;
  timer_begin(ws, v, "scan");
  rc = exclusive_scan(v, v.tiles, v.depth_order, v.offsets, f->slots, v.totals);
  timer_end(ws, v);
  if (rc != 0) return rc;
  GS_HIP_TRY(hipMemcpyAsync(v.host_totals, v.totals, 8, hipMemcpyDeviceToHost, v.stream));
  return 0;
}

int bin_pairs(gs_hip_workspace* ws, View& v, const gs_hip_frame* f, uint64_t pairs) {
  const uint32_t tiles = f->tiles_x * f->tiles_y;
  v.pairs = pairs;
  timer_begin(ws, v, "emit");
  k_fill_ranges<<<div_up((tiles + 1) * 2, 256), 256, 0, v.stream>>>(v.ranges, (tiles + 1) * 2);
  GS_CHECK_LAUNCH("k_fill_ranges");
  if (pairs > 0) {
    k_emit<<<div_up(f->slots, 256), 256, 0, v.stream>>>(
        *f, reinterpret_cast<const float4*>(v.records), reinterpret_cast<const float4*>(v.footprint), v.tiles,
        v.depth_order, v.offsets, v.pair_tiles, v.pair_gaussians, v.pair_capacity, v.totals);
    GS_CHECK_LAUNCH("k_emit");
  }
  timer_end(ws, v);
  if (pairs > 0) {
    timer_begin(ws, v, "tile_sort");
    const int rc = radix_sort(v, v.pair_tiles, v.pair_gaussians, uint32_t(pairs), bits_for(tiles + 1));
    timer_end(ws, v);
    if (rc != 0) return rc;
    timer_begin(ws, v, "ranges");
    k_ranges<<<div_up(pairs, 256), 256, 0, v.stream>>>(v.pair_tiles, uint32_t(pairs),
                                                        reinterpret_cast<uint2*>(v.ranges));
    GS_CHECK_LAUNCH("k_ranges");
    timer_end(ws, v);
  }
  k_tile_order<<<1, 1024, 0, v.stream>>>(reinterpret_cast<const uint2*>(v.ranges), tiles, v.tile_order);
  GS_CHECK_LAUNCH("k_tile_order");
  return 0;
}

}  // namespace gsh

extern "C" {

int gs_hip_project(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "project: null argument");
  if (f->slots > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "project: slots exceed capacity");
  if (f->cam.width != ws->width || f->cam.height != ws->height) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "project: camera size differs from the workspace");
  }
  return gsh::launch_project(ws, gsh::view0(ws), f);
}

int gs_hip_bin(gs_hip_workspace* ws, const gs_hip_frame* f, uint64_t* pairs) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "bin: null argument");
  gsh::View& v = gsh::view0(ws);
  int rc = gsh::bin_count(ws, v, f);
  if (rc != 0) return rc;
  GS_HIP_TRY(hipStreamSynchronize(v.stream));
  const uint64_t p = v.host_totals[0];
  rc = gsh::reserve_pairs(ws, v, p);
  if (rc != 0) return rc;
  rc = gsh::bin_pairs(ws, v, f, p);
  if (pairs != nullptr) *pairs = p;
  return rc;
}

}  // extern "C"
