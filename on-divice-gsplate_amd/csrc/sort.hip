// Binning of every view of a step, one launch per stage: depth sort of the
// slots, exclusive scan of their tile counts in depth order, pair emission,
// stable tile sort of the pairs, per-tile ranges, and the rasterizers' work
// order.  The order this produces is that of sorting 64-bit keys
// (tile << 32) | depth_key with ties broken by the gaussian index: the slots
// are sorted stably by depth key, each emits its pairs at its offset in that
// order, and a stable sort by tile alone then leaves each tile's pairs front
// to back (engine/src/trainer.cpp:665-711, train_emit.comp:3-9).
//
// Counts stay on the device: a sort or scan segment reads its length from
// device memory (the frame's slot count, the pair total) and its launch is
// sized for the buffer's capacity, so no step waits for the host between
// stages.  Pairs beyond the pair buffers are dropped and the step is flagged
// as overflowed (the optimizer then leaves the model alone and the host runs
// the step again with larger buffers, as trainer.cpp:819-836 does).
//
// The radix sort is LSD with 4..8-bit digits; a block owns 4096 items, each
// wave a contiguous 1024 of them.  Per pass: an upsweep counts the block's
// digits (per-wave LDS histograms, one add per group of lanes sharing a
// digit), a scan turns the [digit][block] counts into offsets, and a scatter
// ranks the items stably (wave64 ballots find the lanes holding the same
// digit, rank = popcount of those below plus the wave's running count of
// that digit), places them in LDS in digit order, and writes each digit's
// run out contiguously.
#include <algorithm>

#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;
using gsh::ScanSeg;
using gsh::SortSeg;
using gsh::ViewDev;
using gsh::div_up;
using gsh::kCountingMaxBins;
using gsh::kTileChunk;

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortItems = kSortThreads * kSortRounds;  // 4096 per block
constexpr int kWaves = kSortThreads / 64;

__device__ __forceinline__ uint32_t seg_count(const uint32_t* count, uint32_t count_max) {
  return count != nullptr ? min(*count, count_max) : count_max;
}

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

// Histogram of block (blockIdx.x) of segment blockIdx.y; blocks past the
// segment's length write zeros.  Each wave counts a contiguous 1024 items
// into its own LDS histogram with one LDS atomic per item (a count does not
// depend on the order of the additions), or one per round when the round's
// items share their digit (the top digit of a depth key mostly does: 64
// lanes on one counter would take 64 passes).
__global__ void __launch_bounds__(kSortThreads) k_radix_upsweep(const SortSeg* __restrict__ segs, int shift,
                                                                int nbits, int from_alt) {
  // Four counters per digit and wave, one per lane mod 4 (rows of 257
  // words: a digit's four counters sit in four banks), so lanes of one
  // atomic instruction that share a digit mostly add to different words.
  constexpr int kSplit = 4, kRow = 257;
  __shared__ uint32_t h[kWaves][kSplit * kRow];
  const SortSeg& g = segs[blockIdx.y];
  if (blockIdx.x >= g.blocks) return;  // a smaller segment of the launch
  const uint32_t n = seg_count(g.count, g.count_max);
  const uint32_t radix = 1u << nbits, mask = radix - 1u;
  if (blockIdx.x * kSortItems >= n) {  // past the segment's items: zero counts
    for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) gmem(g.hist)[size_t(d) * g.blocks + blockIdx.x] = 0u;
    return;
  }
  const GS_GLOBAL uint32_t* __restrict__ keys = gmem(from_alt ? g.keys_alt : g.keys);
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  for (int d = threadIdx.x; d < kWaves * kSplit * kRow; d += kSortThreads) (&h[0][0])[d] = 0u;
  __syncthreads();
  uint32_t* hw = &h[wave][(lane & (kSplit - 1)) * kRow];
  const uint32_t wbase = blockIdx.x * kSortItems + wave * (kSortItems / kWaves);
  uint32_t kv[kSortRounds];  // every round's key requested before any is counted
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t idx = wbase + r * 64 + lane;
    kv[r] = idx < n ? keys[idx] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t digit = (kv[r] >> shift) & mask;
    const uint64_t active = __ballot(valid);
    if (active == 0ull) continue;
    const uint32_t first = __builtin_amdgcn_readfirstlane(valid ? digit : 0xffffffffu);
    if (__ballot(valid && digit == first) == active) {
      if (lane == __ffsll(static_cast<long long>(active)) - 1) atomicAdd(&hw[first], uint32_t(__popcll(active)));
    } else if (valid) {
      atomicAdd(&hw[digit], 1u);
    }
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) {
    uint32_t s = 0;
    for (int w = 0; w < kWaves; ++w)
      for (int c = 0; c < kSplit; ++c) s += h[w][c * kRow + d];
    gmem(g.hist)[size_t(d) * g.blocks + blockIdx.x] = s;
  }
}

// One block per (digit, segment): exclusive scan of that digit's per-block
// counts (the blocks the pass launched, up to the segment's scratch), and the
// digit's total.
__global__ void __launch_bounds__(256) k_radix_scan(const SortSeg* __restrict__ segs, uint32_t launched) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_carry;
  const SortSeg& g = segs[blockIdx.y];
  const uint32_t blocks = min(launched, g.blocks);
  uint32_t* row = g.hist + size_t(blockIdx.x) * g.blocks;
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
  if (threadIdx.x == 0) g.digit_totals[blockIdx.x] = s_carry;
}

// Exclusive scan of one value per thread of a 256-thread block into out[thread].
__device__ __forceinline__ void block_scan_256(uint32_t v, uint32_t* s_tmp, uint32_t* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t below = incl - v;
  for (int w = 0; w < wave; ++w) below += s_tmp[w];
  out[threadIdx.x] = below;
  __syncthreads();
}

// Stable scatter of one block: wave w ranks its contiguous 1024 items
// (items base + 1024 w + 64 r + lane, so (w, r, lane) is the input order)
// against a running per-wave digit count in LDS; one block-level combine
// turns (wave, digit) counts into offsets; items go to LDS in digit order and
// each digit's run leaves contiguously.
__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const SortSeg* __restrict__ segs, int shift,
                                                                int nbits, int from_alt) {
  __shared__ uint32_t s_keys[kSortItems];
  __shared__ uint32_t s_vals[kSortItems];
  __shared__ uint32_t s_wave[kWaves][256];  // per-wave digit counts, then each wave's offset inside the digit
  __shared__ uint32_t s_local[256];         // block-local start of each digit
  __shared__ uint32_t s_global[256];        // where the block's run of each digit starts in the output
  __shared__ uint32_t s_tmp[kWaves];
  const SortSeg& g = segs[blockIdx.y];
  const uint32_t n = seg_count(g.count, g.count_max);
  const uint32_t base = blockIdx.x * kSortItems;
  if (base >= n || blockIdx.x >= g.blocks) return;
  const GS_GLOBAL uint32_t* __restrict__ keys_in = gmem(from_alt ? g.keys_alt : g.keys);
  const GS_GLOBAL uint32_t* __restrict__ vals_in = gmem(from_alt ? g.vals_alt : g.vals);
  GS_GLOBAL uint32_t* __restrict__ keys_out = gmem(from_alt ? g.keys : g.keys_alt);
  GS_GLOBAL uint32_t* __restrict__ vals_out = gmem(from_alt ? g.vals : g.vals_alt);
  const uint32_t radix = 1u << nbits, mask = radix - 1u;
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  // Output start of each digit: the digits below it over the whole segment
  // (a block scan of the totals) plus this digit's items in earlier blocks.
  block_scan_256(threadIdx.x < radix ? gmem(g.digit_totals)[threadIdx.x] : 0u, s_tmp, s_global);
  if (threadIdx.x < radix) s_global[threadIdx.x] += gmem(g.hist)[size_t(threadIdx.x) * g.blocks + blockIdx.x];
  for (int d = threadIdx.x; d < kWaves * 256; d += kSortThreads) (&s_wave[0][0])[d] = 0u;
  __syncthreads();
  const uint32_t wbase = base + wave * (kSortItems / kWaves);
  uint32_t key[kSortRounds], val[kSortRounds], wpos[kSortRounds];
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t idx = wbase + r * 64 + lane;
    key[r] = idx < n ? keys_in[idx] : 0u;
    val[r] = idx < n ? vals_in[idx] : 0u;
  }
  // Wave-local stable ranks: earlier items of the same digit in this wave's chunk.
  const uint64_t below_me = lanemask_lt();
  for (int r = 0; r < kSortRounds; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t digit = (key[r] >> shift) & mask;
    const uint64_t peers = match_digit(digit, nbits, __ballot(valid));
    const uint32_t before = valid ? s_wave[wave][digit] : 0u;  // every peer reads, then the first one writes
    if (valid && (peers & below_me) == 0ull) s_wave[wave][digit] = before + __popcll(peers);
    wpos[r] = before + __popcll(peers & below_me);
  }
  __syncthreads();
  // Per digit: the waves' counts -> each wave's offset inside the digit, and the block's count.
  uint32_t total = 0u;
  if (threadIdx.x < radix) {
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_wave[w][threadIdx.x];
      s_wave[w][threadIdx.x] = total;
      total += c;
    }
  }
  block_scan_256(total, s_tmp, s_local);
  for (int r = 0; r < kSortRounds; ++r) {
    if (wbase + r * 64 + lane < n) {
      const uint32_t digit = (key[r] >> shift) & mask;
      const uint32_t pos = s_local[digit] + s_wave[wave][digit] + wpos[r];
      s_keys[pos] = key[r];
      s_vals[pos] = val[r];
    }
  }
  __syncthreads();
  // Each digit's run leaves contiguously: consecutive threads, consecutive addresses.
  const uint32_t items = min(uint32_t(kSortItems), n - base);
  for (uint32_t i = threadIdx.x; i < items; i += kSortThreads) {
    const uint32_t k = s_keys[i];
    const uint32_t digit = (k >> shift) & mask;
    const uint32_t pos = s_global[digit] + (i - s_local[digit]);
    keys_out[pos] = k;
    vals_out[pos] = s_vals[i];
  }
}

// After an odd number of passes: the result back into (keys, vals).
__global__ void __launch_bounds__(256) k_radix_copy_back(const SortSeg* __restrict__ segs) {
  const SortSeg& g = segs[blockIdx.y];
  const uint32_t n = seg_count(g.count, g.count_max);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    g.keys[i] = g.keys_alt[i];
    g.vals[i] = g.vals_alt[i];
  }
}

// ---- exclusive scan ----------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanRounds = 16;
constexpr int kScanItems = kScanThreads * kScanRounds;

__device__ __forceinline__ uint32_t load_value(const ScanSeg& s, uint32_t idx) {
  return s.gather != nullptr ? gmem(s.values)[gmem(s.gather)[idx]] : gmem(s.values)[idx] >> s.shift;
}

// A thread's kScanRounds values of a scan block, all requested before any is
// used (zero past the count).
__device__ __forceinline__ void load_rounds(const ScanSeg& g, uint32_t base, uint32_t count,
                                            uint32_t (&v)[kScanRounds]) {
  const GS_GLOBAL uint32_t* values = gmem(g.values);
  if (g.gather != nullptr) {  // the indices first, then the values they name
    const GS_GLOBAL uint32_t* gather = gmem(g.gather);
    uint32_t at[kScanRounds];
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
      const uint32_t idx = base + r * kScanThreads + threadIdx.x;
      at[r] = idx < count ? gather[idx] : 0xffffffffu;
    }
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) v[r] = at[r] != 0xffffffffu ? values[at[r]] : 0u;
  } else {
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
      const uint32_t idx = base + r * kScanThreads + threadIdx.x;
      v[r] = idx < count ? values[idx] >> g.shift : 0u;
    }
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const ScanSeg* __restrict__ segs) {
  __shared__ uint32_t s[kScanThreads / 64];
  const ScanSeg& g = segs[blockIdx.y];
  if (blockIdx.x >= g.blocks) return;  // a smaller segment of the launch
  const uint32_t count = seg_count(g.count, g.count_max);
  const uint32_t base = blockIdx.x * kScanItems;
  uint32_t vals[kScanRounds];
  load_rounds(g, base, count, vals);
  uint32_t sum = 0;
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) sum += vals[r];
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) g.block_sums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ void __launch_bounds__(1024) k_scan_blocks(const ScanSeg* __restrict__ segs, uint32_t blocks) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_carry;
  const ScanSeg& g = segs[blockIdx.y];
  uint32_t* block_sums = g.block_sums;
  blocks = min(blocks, g.blocks);
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
  if (threadIdx.x == 0) *g.total = s_carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(const ScanSeg* __restrict__ segs) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  __shared__ uint32_t s_carry;
  const ScanSeg& g = segs[blockIdx.y];
  const uint32_t count = seg_count(g.count, g.count_max);
  const uint32_t base = blockIdx.x * kScanItems;
  if (base >= count || blockIdx.x >= g.blocks) return;
  if (threadIdx.x == 0) s_carry = g.block_sums[blockIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t vals[kScanRounds];
  load_rounds(g, base, count, vals);
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) {
    const uint32_t idx = base + r * kScanThreads + threadIdx.x;
    const uint32_t v = vals[r];
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    if (idx < count) gmem(g.out)[idx] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
}

// ---- pair emission and ranges ------------------------------------------------
// One thread per slot in depth order: writes the slot's pairs (tile, slot) at
// its offset, row by row over the same footprint the projection counted
// (train_emit.comp:26-64).  A slot with more than kWaveEmit tiles is written
// by its whole wave instead: the wave walks its tile rows together and the
// lanes write one row's tiles side by side, so one large splat neither holds
// its wave for hundreds of iterations nor scatters single-lane stores.  Pairs
// past the buffer are dropped (the step is then an overflow); a count that
// differs (it cannot: same code, same inputs) is padded with the unused tile
// `tiles` and counted in totals[1].
constexpr uint32_t kWaveEmit = 32;
// Pairs of a wave's small splats staged per pass in LDS, then stored
// contiguously (lane l writes positions l, l + 64, ...): the wave's slots are
// consecutive in depth order, so their pairs form one run [lo, hi).
constexpr uint32_t kEmitWindow = 512;
constexpr uint32_t kNoTile = 0xffffffffu;

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = (uint64_t(__shfl_xor(uint32_t(v >> 32), o, 64)) << 32) | __shfl_xor(uint32_t(v), o, 64);
    v = min(v, w);
  }
  return uniform_u64(v);
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = (uint64_t(__shfl_xor(uint32_t(v >> 32), o, 64)) << 32) | __shfl_xor(uint32_t(v), o, 64);
    v = max(v, w);
  }
  return uniform_u64(v);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(256) k_emit(const ViewDev* __restrict__ views) {
  __shared__ uint32_t s_tile[4][kEmitWindow];
  __shared__ uint32_t s_gauss[4][kEmitWindow];
  const ViewDev& v = views[blockIdx.y];
  const gs_hip_frame& f = v.f;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
  uint32_t i = 0u, count = 0u;
  float4 fp = make_float4(0, 0, 0, 0), r0 = fp;
  if (k < f.slots) {
    i = gmem(v.depth_order)[k];
    // One 32 B line per slot: the depth order makes these reads random.
    fp = gload(v.footprint + size_t(i) * 2u);
    r0 = gload(v.footprint + size_t(i) * 2u + 1u);  // mean x, mean y, tiles
    count = __float_as_uint(r0.z);
  }
  const uint64_t p0 = count > 0u ? uint64_t(gmem(v.offsets)[k]) : 0ull;
  const uint64_t cap = v.pair_capacity;
  const uint32_t pad_tile = v.bin_count;
  const int side_log2 = 4 + int(v.bin_shift);
  GS_GLOBAL uint32_t* const out_tiles = gmem(v.emit_tiles);
  GS_GLOBAL uint32_t* const out_gauss = gmem(v.emit_gaussians);
  const bool small = count > 0u && count <= kWaveEmit;
  const uint64_t end = min(p0 + count, cap);
  int row_first = 0, row_last = -1;
  if (small) footprint_rows(r0.y, fp.z, fp.w, f.cam.height, row_first, row_last);
  const RowSpan span = row_span(r0.x, r0.y, fp.x, fp.y, fp.z, fp.w);
  // The run of the wave's small-splat pairs (wave-uniform); windows skip the
  // stretches that only large splats fill.
  const uint64_t hi = wave_max_u64(small ? end : 0ull);
  const uint64_t lo = wave_min_u64(small && p0 < end ? p0 : ~0ull);
  for (uint64_t win = lo; win < hi;) {
    for (uint32_t e = uint32_t(lane); e < kEmitWindow; e += 64u) s_tile[wave][e] = kNoTile;
    wave_sync();
    if (small) {
      // The slot's tiles row by row, as the projection counted them (train_emit.comp:26-64);
      // the ones in this window go to LDS.
      uint64_t p = p0;
      uint32_t wanted = 0;
      if (row_first <= row_last) {
        for (uint32_t ty = uint32_t(row_first) >> side_log2; ty <= (uint32_t(row_last) >> side_log2); ++ty) {
          uint32_t x0, x1;
          span_row_tiles(span, row_first, row_last, ty, f.cam.width, x0, x1, side_log2);
          wanted += x1 - x0;
          for (uint32_t x = x0; x < x1 && p < end; ++x, ++p) {
            if (p >= win && p < win + kEmitWindow) {
              s_tile[wave][p - win] = ty * v.bins_x + x;
              s_gauss[wave][p - win] = i;
            }
          }
        }
      }
      if (wanted != count) {
        if (win == lo) atomicAdd(&v.totals[1], 1u);
        for (; p < end; ++p) {
          if (p >= win && p < win + kEmitWindow) {
            s_tile[wave][p - win] = pad_tile;
            s_gauss[wave][p - win] = i;
          }
        }
      }
    }
    wave_sync();
    for (uint32_t e = uint32_t(lane); e < kEmitWindow && win + e < hi; e += 64u) {
      const uint32_t t = s_tile[wave][e];
      if (t != kNoTile) {  // positions of the wave's large splats are written below
        out_tiles[win + e] = t;
        out_gauss[win + e] = s_gauss[wave][e];
      }
    }
    wave_sync();
    const uint64_t next = win + kEmitWindow;
    win = wave_min_u64(small && end > next ? max(p0, next) : ~0ull);
  }
  // Large splats, one at a time per wave: every lane takes part.
  uint64_t big = __ballot(count > kWaveEmit);
  while (big != 0ull) {
    const int src = __builtin_ctzll(big);
    big &= big - 1ull;
    const uint32_t bi = __shfl(i, src, 64), bcount = __shfl(count, src, 64);
    const float4 bfp = make_float4(__shfl(fp.x, src, 64), __shfl(fp.y, src, 64), __shfl(fp.z, src, 64),
                                   __shfl(fp.w, src, 64));
    const float bmx = __shfl(r0.x, src, 64), bmy = __shfl(r0.y, src, 64);
    const uint64_t start = (uint64_t(__shfl(uint32_t(p0 >> 32), src, 64)) << 32) | __shfl(uint32_t(p0), src, 64);
    const uint64_t end = min(start + bcount, cap);
    uint64_t p = start;
    uint32_t wanted = 0;
    int row_first, row_last;
    footprint_rows(bmy, bfp.z, bfp.w, f.cam.height, row_first, row_last);
    if (row_first <= row_last) {
      const RowSpan bspan = row_span(bmx, bmy, bfp.x, bfp.y, bfp.z, bfp.w);
      for (uint32_t ty = uint32_t(row_first) >> side_log2; ty <= (uint32_t(row_last) >> side_log2); ++ty) {
        uint32_t x0, x1;
        span_row_tiles(bspan, row_first, row_last, ty, f.cam.width, x0, x1, side_log2);
        wanted += x1 - x0;
        for (uint32_t x = x0 + uint32_t(lane); x < x1; x += 64u) {
          const uint64_t q = p + (x - x0);
          if (q < end) {
            out_tiles[q] = ty * v.bins_x + x;
            out_gauss[q] = bi;
          }
        }
        p += x1 - x0;
      }
    }
    if (wanted != bcount) {
      if (lane == 0) atomicAdd(&v.totals[1], 1u);
      for (uint64_t q = min(p, end) + uint64_t(lane); q < end; q += 64u) {
        out_tiles[q] = pad_tile;
        out_gauss[q] = bi;
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_clear_ranges(const ViewDev* __restrict__ views) {
  const ViewDev& v = views[blockIdx.y];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= v.tile_count) v.ranges[i] = make_uint2(0u, 0u);
  if (i == 0u) v.totals[1] = 0u;
}

// Tile boundaries of the sorted pairs (grid-stride: the grid is bounded, the
// count is on the device).
__global__ void __launch_bounds__(256) k_ranges(const ViewDev* __restrict__ views) {
  const ViewDev& v = views[blockIdx.y];
  const uint32_t count = uint32_t(std::min<uint64_t>(v.totals[0], v.pair_capacity));
  const uint32_t* pt_flat = v.pair_tiles;
  const GS_GLOBAL uint32_t* pt = gmem(pt_flat);
  GS_GLOBAL uint32_t* rg = gmem(reinterpret_cast<uint32_t*>(v.ranges));  // [tile][2]: first, end
  // Four consecutive pairs per thread and iteration (one 16 B load; the
  // buffer is 256 B aligned), plus the pair before them.
  const uint32_t quads = (count + 3u) / 4u;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
    const uint32_t s0 = q * 4u;
    uint32_t t[4];
    if (s0 + 4u <= count) {
      const uint4 w = gload(reinterpret_cast<const uint4*>(pt_flat + s0));
      t[0] = w.x, t[1] = w.y, t[2] = w.z, t[3] = w.w;
    } else {
      for (int i = 0; i < 4; ++i) t[i] = s0 + i < count ? pt[s0 + i] : 0u;
    }
    uint32_t prev = s0 > 0u ? pt[s0 - 1u] : 0u;
    for (int i = 0; i < 4; ++i) {
      const uint32_t s = s0 + uint32_t(i);
      if (s >= count) break;
      const uint32_t tile = t[i];
      if (s == 0u) {
        rg[size_t(tile) * 2u] = 0u;
      } else if (prev != tile) {
        rg[size_t(prev) * 2u + 1u] = s;
        rg[size_t(tile) * 2u] = s;
      }
      if (s == count - 1u) rg[size_t(tile) * 2u + 1u] = count;
      prev = tile;
    }
  }
}

// The rasterizers' work: every (view, tile) of the step, as view << 16 | tile,
// in 2x2 quads of neighbouring tiles (four consecutive items, kNoWork for a
// member outside the grid), the quads bucketed by floor(log2(their longest
// list)), longest first, empty quads last (longest-processing-time-first
// scheduling within a factor of two, across all views; the order within a
// bucket is whatever the atomics give).  The rasterizers hand a quad's four
// items to workgroups on one XCD, so the records the neighbours share (a
// whole bin's list with pair bins wider than a tile) are fetched from HBM
// once into that XCD's L2.  Two launches, a thread per quad:
//   k_work_count: each bucket's quads over the step (per-block LDS counts,
//                 one global atomic per block and bucket); its first block
//                 also clears the persistent rasterizers' work counters and
//                 sets the step's pair-overflow flag;
//   k_work_place: each block takes its run of each bucket (the bucket's
//                 start, the exclusive prefix of the counts, plus the runs
//                 taken before it) and writes its quads' items there; the
//                 last block to finish clears the buckets' counters for the
//                 next step (they are zero from the workspace's creation on).
__device__ __forceinline__ uint32_t quad_bucket(const ViewDev& view, uint32_t q) {
  const uint32_t tx = view.f.tiles_x, ty = view.f.tiles_y, qx = gsh::quads_x(tx);
  const uint32_t x0 = (q % qx) * 2u, y0 = (q / qx) * 2u;
  uint32_t longest = 0u;
  for (uint32_t m = 0; m < 4u; ++m) {
    const uint32_t x = x0 + (m & 1u), y = y0 + (m >> 1);
    if (x < tx && y < ty) {
      const uint2 r = gload(view.ranges + gsh::bin_of_tile(view, y * tx + x));
      longest = max(longest, r.y - r.x);
    }
  }
  return __clz(longest);
}

// The first block's first wave: the persistent rasterizers' work counters,
// and whether any view of the step found more pairs than its buffers hold
// (the optimizer then leaves every model alone: one flag instead of every
// view's count read by every optimizer thread).
__device__ __forceinline__ void clear_step_counters(const ViewDev* __restrict__ views, uint32_t count,
                                                    uint32_t* __restrict__ counters) {
  if (threadIdx.x < 4) counters[threadIdx.x] = 0u;
  bool over = false;
  for (uint32_t v = threadIdx.x; v < count; v += 64u) {
    over = over || uint64_t(views[v].totals[0]) > views[v].pair_capacity;
  }
  const uint64_t any = __ballot(over);
  if (threadIdx.x == 0) counters[gsh::kStepOverflowCounter] = any != 0ull ? 1u : 0u;
}

// This thread's quad's bucket and its rank among the block's quads of that
// bucket, counted in s_count (lanes of a wave that share a bucket add their
// count once: an LDS atomic from 64 lanes on one counter would take 64
// passes).
__device__ __forceinline__ uint32_t count_quad(uint32_t b, uint32_t* s_count) {
  const uint64_t peers = match_digit(b, 6, __ballot(true));
  const uint64_t below_me = lanemask_lt();
  const int leader = __ffsll(static_cast<long long>(peers)) - 1;
  uint32_t first = 0u;
  if ((peers & below_me) == 0ull) first = atomicAdd(&s_count[b], uint32_t(__popcll(peers)));
  first = __shfl(first, leader, 64);
  return first + uint32_t(__popcll(peers & below_me));
}

__global__ void __launch_bounds__(256) k_work_count(const ViewDev* __restrict__ views, uint32_t count,
                                                    uint32_t* __restrict__ counters) {
  __shared__ uint32_t s_count[gsh::kWorkBuckets];
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 64) clear_step_counters(views, count, counters);
  const ViewDev& view = views[blockIdx.y];
  const uint32_t quads = gsh::quad_items(view.f.tiles_x, view.f.tiles_y) / 4u;
  if (blockIdx.x * 256u >= quads) return;  // the whole block
  if (threadIdx.x < gsh::kWorkBuckets) s_count[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q < quads) count_quad(quad_bucket(view, q), s_count);
  __syncthreads();
  if (threadIdx.x < gsh::kWorkBuckets && s_count[threadIdx.x] != 0u) {
    atomicAdd(counters + gsh::kWorkBucketCounters + threadIdx.x, s_count[threadIdx.x]);
  }
}

__global__ void __launch_bounds__(256) k_work_place(const ViewDev* __restrict__ views, uint32_t* __restrict__ work,
                                                    uint32_t items, uint32_t* __restrict__ counters) {
  __shared__ uint32_t s_count[gsh::kWorkBuckets];
  __shared__ uint32_t s_base[gsh::kWorkBuckets];
  __shared__ uint32_t s_last;
  const uint32_t v = blockIdx.y;
  const ViewDev& view = views[v];
  const uint32_t tx = view.f.tiles_x, ty = view.f.tiles_y, qx = gsh::quads_x(tx);
  const uint32_t quads = gsh::quad_items(tx, ty) / 4u;
  // Blocks past a view's quads place nothing but still count as finished.
  if (threadIdx.x < gsh::kWorkBuckets) s_count[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  uint32_t b = 0u, rank = 0u;
  if (q < quads) {
    b = quad_bucket(view, q);
    rank = count_quad(b, s_count);
  }
  __syncthreads();
  if (threadIdx.x < gsh::kWorkBuckets) {
    const uint32_t k = threadIdx.x;
    const GS_GLOBAL uint32_t* totals = gmem(counters) + gsh::kWorkBucketCounters;
    uint32_t start = 0u;
    for (uint32_t j = 0; j < k; ++j) start += totals[j];
    const uint32_t mine = s_count[k];
    s_base[k] = mine != 0u ? start + atomicAdd(counters + gsh::kWorkBucketCounters + gsh::kWorkBuckets + k, mine) : 0u;
  }
  __syncthreads();
  // (A place past the list could only come from counts left by a step that
  // stopped between the two launches; it is dropped, not written.)
  if (q < quads && 4u * (s_base[b] + rank) + 4u <= items) {
    const uint32_t at = 4u * (s_base[b] + rank);
    const uint32_t x0 = (q % qx) * 2u, y0 = (q / qx) * 2u;
    for (uint32_t m = 0; m < 4u; ++m) {
      const uint32_t x = x0 + (m & 1u), y = y0 + (m >> 1);
      work[at + m] = x < tx && y < ty ? (v << gsh::kTileBits) | (y * tx + x) : gsh::kNoWork;
    }
  }
  // Every block has read the counts and taken its runs before it counts
  // itself finished, so the last one may clear them.
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t done = atomicAdd(counters + gsh::kWorkDoneCounter, 1u);
    s_last = done + 1u == gridDim.x * gridDim.y ? 1u : 0u;
  }
  __syncthreads();
  if (s_last != 0u) {
    __threadfence();
    for (uint32_t i = threadIdx.x; i < 2u * gsh::kWorkBuckets; i += 256u) counters[gsh::kWorkBucketCounters + i] = 0u;
    if (threadIdx.x == 0) counters[gsh::kWorkDoneCounter] = 0u;
  }
}

// ---- counting tile sort ------------------------------------------------------
// One stable counting pass over the pairs in emission order instead of LSD
// radix passes: the tile grid (<= kCountingMaxBins bins, the last one for
// slots whose tiles were miscounted) fits one chunk's counters in LDS.
//   k_tile_hist:    per chunk of kTileChunk pairs, its count of each tile
//                   ([col][bin] table, LDS atomics: a count ignores order);
//   k_tile_colscan: per tile, the exclusive prefix over chunks (in place)
//                   and the tile's total;
//   k_tile_starts:  per view, the exclusive prefix over tiles: the ranges;
//   k_tile_scatter: one wave per chunk walks its pairs in order, ranking
//                   each among equal tiles of its round with ballots and
//                   placing it at the tile's start + earlier chunks' count +
//                   the chunk's running count (LDS), so each tile keeps the
//                   emission (depth) order.
// Chunks are dealt to the XCDs in contiguous runs, so one XCD writes one
// contiguous stretch of each tile's range and its L2 merges the lines.

__device__ __forceinline__ uint32_t pair_count(const ViewDev& v) {
  return uint32_t(min(uint64_t(v.totals[0]), v.pair_capacity));
}

__global__ void __launch_bounds__(256) k_tile_hist(const ViewDev* __restrict__ views) {
  extern __shared__ uint32_t s_bins[];
  const ViewDev& v = views[blockIdx.y];
  const uint32_t n = pair_count(v), bins = v.bin_count + 1u;
  const uint32_t col = blockIdx.x;
  if (uint64_t(col) * kTileChunk >= n) return;
  for (uint32_t t = threadIdx.x; t < bins; t += blockDim.x) s_bins[t] = 0u;
  __syncthreads();
  const GS_GLOBAL uint32_t* keys = gmem(v.emit_tiles);
  const uint32_t first = col * kTileChunk, last = uint32_t(min(uint64_t(first) + kTileChunk, uint64_t(n)));
  constexpr int kUnroll = 8;
  for (uint32_t base = first; base < last; base += 256u * kUnroll) {
    uint32_t key[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const uint32_t idx = base + uint32_t(u) * 256u + threadIdx.x;
      key[u] = idx < last ? keys[idx] : 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (key[u] < bins) atomicAdd(&s_bins[key[u]], 1u);
  }
  __syncthreads();
  GS_GLOBAL uint32_t* row = gmem(v.tile_hist) + size_t(col) * bins;
  for (uint32_t t = threadIdx.x; t < bins; t += blockDim.x) row[t] = s_bins[t];
}

// One thread per (view, bin): the chunks' counts of the bin become their
// exclusive prefix; the bin's total goes to ranges[bin].y until the starts
// are known.  The counts are requested in groups of 16 before they are summed.
__global__ void __launch_bounds__(256) k_tile_colscan(const ViewDev* __restrict__ views) {
  const ViewDev& v = views[blockIdx.y];
  const uint32_t bins = v.bin_count + 1u;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bins) return;
  const uint32_t cols = uint32_t((uint64_t(pair_count(v)) + kTileChunk - 1u) / kTileChunk);
  GS_GLOBAL uint32_t* table = gmem(v.tile_hist);
  uint32_t run = 0u;
  constexpr uint32_t kGroup = 16;
  for (uint32_t c0 = 0; c0 < cols; c0 += kGroup) {
    uint32_t x[kGroup];
#pragma unroll
    for (uint32_t k = 0; k < kGroup; ++k) x[k] = c0 + k < cols ? table[size_t(c0 + k) * bins + t] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kGroup; ++k) {
      if (c0 + k < cols) table[size_t(c0 + k) * bins + t] = run;
      run += x[k];
    }
  }
  gmem(reinterpret_cast<uint32_t*>(v.ranges))[size_t(t) * 2u + 1u] = run;
}

// One workgroup per view: the exclusive prefix of the bins' totals; each
// bin's range (first, end), (0, 0) for an empty one.
__global__ void __launch_bounds__(1024) k_tile_starts(const ViewDev* __restrict__ views) {
  __shared__ uint32_t s_wave[16];
  const ViewDev& v = views[blockIdx.y];
  const uint32_t bins = v.bin_count + 1u;
  const uint32_t per = (bins + 1023u) / 1024u;  // <= kCountingMaxBins / 1024
  GS_GLOBAL uint32_t* rg = gmem(reinterpret_cast<uint32_t*>(v.ranges));
  const uint32_t t0 = threadIdx.x * per;
  uint32_t tot[kCountingMaxBins / 1024];
  uint32_t sum = 0u;
#pragma unroll
  for (uint32_t k = 0; k < kCountingMaxBins / 1024; ++k) {
    tot[k] = k < per && t0 + k < bins ? rg[size_t(t0 + k) * 2u + 1u] : 0u;
    sum += tot[k];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t start = incl - sum;
  for (int w = 0; w < wave; ++w) start += s_wave[w];
#pragma unroll
  for (uint32_t k = 0; k < kCountingMaxBins / 1024; ++k) {
    if (k < per && t0 + k < bins) {
      const uint2 r = tot[k] > 0u ? make_uint2(start, start + tot[k]) : make_uint2(0u, 0u);
      rg[size_t(t0 + k) * 2u] = r.x;
      rg[size_t(t0 + k) * 2u + 1u] = r.y;
      start += tot[k];
    }
  }
}

__global__ void __launch_bounds__(64) k_tile_scatter(const ViewDev* __restrict__ views, int key_bits) {
  extern __shared__ uint32_t s_run[];
  // Contiguous runs of (view, chunk) per XCD: workgroups go to the XCDs in
  // turn, and the grid's size is a multiple of 8.
  const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;
  const uint32_t idx = (lin % 8u) * (gridDim.x * gridDim.y / 8u) + lin / 8u;
  const uint32_t view = idx / gridDim.x, col = idx % gridDim.x;
  const ViewDev& v = views[view];
  const uint32_t n = pair_count(v), bins = v.bin_count + 1u;
  if (uint64_t(col) * kTileChunk >= n) return;
  const uint32_t lane = threadIdx.x;
  const GS_GLOBAL uint32_t* prefix = gmem(v.tile_hist) + size_t(col) * bins;
  const GS_GLOBAL uint32_t* rg = gmem(reinterpret_cast<const uint32_t*>(v.ranges));
  for (uint32_t t = lane; t < bins; t += 64u) s_run[t] = rg[size_t(t) * 2u] + prefix[t];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const GS_GLOBAL uint32_t* in_t = gmem(v.emit_tiles);
  const GS_GLOBAL uint32_t* in_g = gmem(v.emit_gaussians);
  GS_GLOBAL uint32_t* out_t = gmem(v.pair_tiles);
  GS_GLOBAL uint32_t* out_g = gmem(v.pair_gaussians);
  const uint64_t below_me = lanemask_lt();
  const uint32_t first = col * kTileChunk, last = uint32_t(min(uint64_t(first) + kTileChunk, uint64_t(n)));
  constexpr int kUnroll = 8;
  for (uint32_t base = first; base < last; base += 64u * kUnroll) {
    uint32_t key[kUnroll], val[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const uint32_t i = base + uint32_t(u) * 64u + lane;
      key[u] = i < last ? in_t[i] : 0u;
      val[u] = i < last ? in_g[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool valid = base + uint32_t(u) * 64u + lane < last && key[u] < bins;
      const uint64_t active = __ballot(valid);
      if (active == 0ull) continue;
      const uint64_t peers = match_digit(key[u], key_bits, active);
      const uint32_t before = valid ? s_run[key[u]] : 0u;  // every peer reads, then the first one writes
      if (valid && (peers & below_me) == 0ull) s_run[key[u]] = before + uint32_t(__popcll(peers));
      if (valid) {
        const uint32_t pos = before + uint32_t(__popcll(peers & below_me));
        out_t[pos] = key[u];
        out_g[pos] = val[u];
      }
    }
  }
}

int counting_tile_sort(gs_hip_workspace* ws, uint32_t count) {
  const uint32_t bins = ws->shape.max_bins + 1u;
  const uint32_t cols = std::max<uint32_t>(div_up(ws->shape.max_pair_capacity, kTileChunk), 1u);
  const size_t lds = size_t(bins) * 4u;
  k_tile_hist<<<dim3(cols, count), 256, lds, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_tile_hist");
  k_tile_colscan<<<dim3(div_up(bins, 256), count), 256, 0, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_tile_colscan");
  k_tile_starts<<<dim3(1, count), 1024, 0, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_tile_starts");
  const uint32_t grid_x = (count % 8u == 0u) ? cols : (cols + 7u) / 8u * 8u;  // grid_x * count divisible by 8
  k_tile_scatter<<<dim3(grid_x, count), 64, lds, ws->stream>>>(ws->dev_views, int(ws->shape.tile_key_bits));
  GS_CHECK_LAUNCH("k_tile_scatter");
  return 0;
}

// Batched LSD sort of `count` segments by key bits [0, bits), `digit` bits per pass.
int radix_sort(gs_hip_workspace* ws, const SortSeg* segs, uint32_t count, uint64_t items_max, int bits, int digit) {
  if (bits <= 0 || count == 0 || items_max == 0) return 0;
  const uint32_t blocks = div_up(items_max, kSortItems);
  const int passes = (bits + digit - 1) / digit;
  const dim3 grid(blocks, count);
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * digit;
    const int nbits = std::min(digit, bits - shift);
    const int from_alt = pass & 1;
    k_radix_upsweep<<<grid, kSortThreads, 0, ws->stream>>>(segs, shift, nbits, from_alt);
    GS_CHECK_LAUNCH("k_radix_upsweep");
    k_radix_scan<<<dim3(1u << nbits, count), 256, 0, ws->stream>>>(segs, blocks);
    GS_CHECK_LAUNCH("k_radix_scan");
    k_radix_scatter<<<grid, kSortThreads, 0, ws->stream>>>(segs, shift, nbits, from_alt);
    GS_CHECK_LAUNCH("k_radix_scatter");
  }
  if (passes & 1) {
    k_radix_copy_back<<<dim3(std::min<uint32_t>(div_up(items_max, 256), 4096), count), 256, 0, ws->stream>>>(segs);
    GS_CHECK_LAUNCH("k_radix_copy_back");
  }
  return 0;
}

int scan_segments(gs_hip_workspace* ws, const ScanSeg* segs, uint32_t count, uint64_t items_max) {
  const uint32_t blocks = std::max<uint32_t>(div_up(items_max, kScanItems), 1);
  k_scan_reduce<<<dim3(blocks, count), kScanThreads, 0, ws->stream>>>(segs);
  GS_CHECK_LAUNCH("k_scan_reduce");
  k_scan_blocks<<<dim3(1, count), 1024, 0, ws->stream>>>(segs, blocks);
  GS_CHECK_LAUNCH("k_scan_blocks");
  k_scan_down<<<dim3(blocks, count), kScanThreads, 0, ws->stream>>>(segs);
  GS_CHECK_LAUNCH("k_scan_down");
  return 0;
}

}  // namespace

namespace gsh {

int exclusive_scan(gs_hip_workspace* ws, View& v, const uint32_t* values, const uint32_t* gather, uint32_t* out,
                   uint32_t count, uint32_t* total_dev) {
  const uint32_t blocks = std::max<uint32_t>(div_up(count, kScanItems), 1);
  if (blocks > v.block_sums_entries) return set_error(GS_HIP_E_CAPACITY, "exclusive_scan: too many blocks");
  const ScanSeg seg{values, gather, out, v.block_sums, total_dev, nullptr, count, v.block_sums_entries};
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  GS_HIP_TRY(hipMemcpy(ws->dev_scan_tmp, &seg, sizeof(seg), hipMemcpyHostToDevice));
  return scan_segments(ws, ws->dev_scan_tmp, 1, count);
}

int launch_depth_sort_and_scan(gs_hip_workspace* ws, uint32_t count, bool depth16) {
  const uint32_t items = ws->shape.max_capacity;
  if (items == 0) return 0;
  timer_begin(ws, "depth_sort");
  int rc = radix_sort(ws, ws->dev_depth_sort, count, items, depth16 ? 16 : 32, int(ws->tuning.depth_digit_bits));
  timer_end(ws);
  if (rc != 0) return rc;
  timer_begin(ws, "scan");
  rc = scan_segments(ws, ws->dev_tile_scan, count, items);
  timer_end(ws);
  return rc;
}

int launch_clear_ranges(gs_hip_workspace* ws, uint32_t count) {
  k_clear_ranges<<<dim3(div_up(ws->shape.max_tiles + 1, 256), count), 256, 0, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_clear_ranges");
  return 0;
}

int launch_tile_sort_and_ranges(gs_hip_workspace* ws, uint32_t count) {
  const uint64_t cap = ws->shape.max_pair_capacity;
  timer_begin(ws, "tile_sort");
  int rc = 0;
  if (ws->shape.counting_tile_sort) {
    rc = counting_tile_sort(ws, count);  // also writes the ranges
  } else {
    const int digit = ws->tuning.tile_digit_bits != 0u ? int(ws->tuning.tile_digit_bits) : 8;
    rc = radix_sort(ws, ws->dev_tile_sort, count, cap, int(ws->shape.tile_key_bits), digit);
  }
  timer_end(ws);
  if (rc != 0) return rc;
  timer_begin(ws, "ranges");
  if (!ws->shape.counting_tile_sort) {
    k_ranges<<<dim3(std::min<uint32_t>(div_up(cap, 256), 2048), count), 256, 0, ws->stream>>>(ws->dev_views);
    GS_CHECK_LAUNCH("k_ranges");
  }
  const dim3 quad_grid(std::max<uint32_t>(div_up(ws->shape.max_quads, 256), 1), count);
  k_work_count<<<quad_grid, 256, 0, ws->stream>>>(ws->dev_views, count, ws->work_counters);
  GS_CHECK_LAUNCH("k_work_count");
  k_work_place<<<quad_grid, 256, 0, ws->stream>>>(ws->dev_views, ws->work, ws->shape.work_items, ws->work_counters);
  GS_CHECK_LAUNCH("k_work_place");
  timer_end(ws);
  return 0;
}

int launch_emit_and_tile_sort(gs_hip_workspace* ws, uint32_t count) {
  timer_begin(ws, "emit");
  int rc = launch_clear_ranges(ws, count);
  if (rc == 0 && ws->shape.max_capacity > 0) {
    k_emit<<<dim3(div_up(ws->shape.max_capacity, 256), count), 256, 0, ws->stream>>>(ws->dev_views);
    GS_CHECK_LAUNCH("k_emit");
  }
  timer_end(ws);
  return rc != 0 ? rc : launch_tile_sort_and_ranges(ws, count);
}

}  // namespace gsh

extern "C" {

int gs_hip_project(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "project: null argument");
  if (f->slots > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "project: slots exceed capacity");
  if (f->cam.width != ws->width || f->cam.height != ws->height) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "project: camera size differs from the workspace");
  }
  const int rc = gsh::upload_table(ws, f, 1);
  return rc != 0 ? rc : gsh::launch_project(ws, 1);
}

int gs_hip_bin(gs_hip_workspace* ws, const gs_hip_frame* f, uint64_t* pairs) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "bin: null argument");
  int rc = gsh::upload_table(ws, f, 1);
  if (rc == 0) rc = gsh::launch_depth_sort_and_scan(ws, 1, (f->flags & GS_HIP_FLAG_DEPTH16) != 0u);
  if (rc != 0) return rc;
  gsh::View& v = gsh::view0(ws);
  uint32_t p = 0;
  GS_HIP_TRY(hipMemcpyAsync(&p, v.totals, 4, hipMemcpyDeviceToHost, ws->stream));
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  if (p > v.pair_capacity) {
    rc = gsh::reserve_pairs(ws, p);
    if (rc == 0) rc = gsh::upload_table(ws, f, 1);
    if (rc != 0) return rc;
  }
  v.pairs = p;
  ws->max_pairs_seen = std::max<uint64_t>(ws->max_pairs_seen, p);
  rc = gsh::launch_emit_and_tile_sort(ws, 1);
  if (pairs != nullptr) *pairs = p;
  return rc;
}

}  // extern "C"
