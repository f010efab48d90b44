// The stored-state backward's bucket slots (gsh::bucket_slot and
// gsh::bucket_slots_for in csrc/gs_internal.h, the code the kernels compile),
// on the host: for random partitions of a pair list into bins (lengths from 0
// to several thousand, many empty bins, long single bins) at bin shifts 0, 1
// and 2, every bucket k >= 1 of every raster tile of every bin that its list
// can need (positions 64 k < the bin's length) gets a slot of its own, and
// every slot lies below the view's reserve for that many pairs.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "gs_internal.h"
#include "test_util.h"

namespace {

void check_layout(uint32_t tiles_x, uint32_t tiles_y, uint32_t shift, std::mt19937& rng, int& failures) {
  const uint32_t side = 1u << shift;
  const uint32_t bins_x = (tiles_x + side - 1) / side, bins_y = (tiles_y + side - 1) / side;
  const uint32_t bins = bins_x * bins_y;
  std::vector<uint32_t> len(bins);
  std::uniform_int_distribution<int> kind(0, 9);
  std::uniform_int_distribution<uint32_t> small(0, 200), large(0, 6000);
  uint64_t pairs = 0;
  for (uint32_t b = 0; b < bins; ++b) {
    const int k = kind(rng);
    len[b] = k < 4 ? 0u : (k < 8 ? small(rng) : large(rng));
    pairs += len[b];
  }
  const uint64_t reserve = gsh::bucket_slots_for(pairs, bins, shift);
  std::vector<uint8_t> used(reserve, 0);
  uint32_t start = 0;
  uint64_t buckets = 0;
  for (uint32_t b = 0; b < bins; ++b) {
    const uint32_t bx = b % bins_x, by = b / bins_x;
    for (uint32_t sy = 0; sy < side; ++sy) {
      for (uint32_t sx = 0; sx < side; ++sx) {
        const uint32_t tx = bx * side + sx, ty = by * side + sy;
        if (tx >= tiles_x || ty >= tiles_y) continue;
        const uint32_t tile = ty * tiles_x + tx;
        for (uint32_t k = 1; uint64_t(k) * gsh::kBucketSplats < len[b]; ++k) {
          const uint64_t slot = gsh::bucket_slot(start, b, k, tile, tiles_x, shift);
          ++buckets;
          if (slot >= reserve) {
            std::printf("FAIL shift %u bin %u tile %u k %u: slot %llu past the reserve %llu\n", shift, b, tile, k,
                        (unsigned long long)slot, (unsigned long long)reserve);
            ++failures;
            return;
          }
          if (used[slot]) {
            std::printf("FAIL shift %u bin %u tile %u k %u: slot %llu taken twice\n", shift, b, tile, k,
                        (unsigned long long)slot);
            ++failures;
            return;
          }
          used[slot] = 1;
        }
      }
    }
    start += len[b];
  }
  if (buckets == 0 && pairs > gsh::kBucketSplats * 4) {
    std::printf("FAIL shift %u: no bucket checked\n", shift);
    ++failures;
  }
}

}  // namespace

int main() {
  std::mt19937 rng(20261204);
  int failures = 0;
  const uint32_t grids[][2] = {{1, 1}, {3, 2}, {80, 45}, {120, 68}, {41, 23}, {7, 130}};
  for (const auto& g : grids) {
    for (uint32_t shift = 0; shift <= 2; ++shift) {
      for (int rep = 0; rep < 40; ++rep) check_layout(g[0], g[1], shift, rng, failures);
    }
  }
  // A single bin holding every pair: its buckets reach the end of the reserve.
  {
    const uint64_t pairs = 1u << 20;
    const uint64_t reserve = gsh::bucket_slots_for(pairs, 1, 0);
    const uint64_t last = gsh::bucket_slot(0, 0, uint32_t(pairs / gsh::kBucketSplats) - 1u, 0, 1, 0);
    EXPECT(last < reserve, "the last bucket of one long list stays in the reserve");
  }
  if (failures != 0) {
    std::printf("%d layout failures\n", failures);
    return 1;
  }
  std::printf("bucket slots: every bucket its own slot within the reserve\n");
  return test::failures() == 0 ? 0 : 1;
}
