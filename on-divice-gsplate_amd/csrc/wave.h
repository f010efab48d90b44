// Wave64 reductions for gfx950: DPP moves inside each 16-lane row and the
// gfx950 permlane16/permlane32 swaps across rows.
#pragma once

#include <hip/hip_runtime.h>

namespace gsd {

template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xf, 0xf, false));
}

// (a, b) -> one register: lanes 0-31 hold a summed over lane pairs (i, i+32)
// and lanes 32-63 hold b summed over the same pairs, or the other way round
// (which half gets which is fixed by the instruction, and the lane map of
// Reduce9 below is measured, not assumed).
__device__ __forceinline__ float swap32_add(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  return __builtin_bit_cast(float, unsigned(r[0])) + __builtin_bit_cast(float, unsigned(r[1]));
}

// The same across 16-lane rows: pairs (i, i+16) inside each 32-lane half.
__device__ __forceinline__ float swap16_add(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  return __builtin_bit_cast(float, unsigned(r[0])) + __builtin_bit_cast(float, unsigned(r[1]));
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xb1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f32<0x4e>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f32<0x124>(v);  // row_ror 4
  v += dpp_f32<0x128>(v);  // row_ror 8: every lane holds its row's sum
  v = swap16_add(v, v);
  return swap32_add(v, v);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o, 64);
    v = t > v ? t : v;
  }
  return v;
}

// Nine wave sums at once by transposition: each exchange step pairs two
// registers, keeps one half of the lanes' partial sums of the first and the
// other half's of the second, and so halves the registers in flight
// (9 -> 5 -> 3 -> 2 -> 1), after which a quad holds one value's total.
// About 27 instructions instead of nine separate 6-step reductions.  The
// lane that ends with value k is found once per kernel by reducing the
// probe values k + 1 (slot()).
struct Reduce9 {
  bool hi8, hi4;  // lane bits 3 and 2: which side of the row_ror 8 / half-mirror exchanges

  __device__ __forceinline__ Reduce9() {
    const int lane = int(threadIdx.x & 63u);
    hi8 = (lane & 8) != 0;
    hi4 = (lane & 4) != 0;
  }

  // a in lanes with !hi, b in lanes with hi, each summed with the partner lane.
  template <int kCtrl>
  __device__ __forceinline__ float pair(float a, float b, bool hi) const {
    const float keep = hi ? b : a;
    const float send = hi ? a : b;
    return keep + dpp_f32<kCtrl>(send);
  }

  __device__ __forceinline__ float reduce(float g0, float g1, float g2, float g3, float g4, float g5, float g6, float g7,
                                          float g8) const {
    const float w0 = swap32_add(g0, g1), w1 = swap32_add(g2, g3), w2 = swap32_add(g4, g5), w3 = swap32_add(g6, g7);
    const float w4 = swap32_add(g8, 0.0f);
    const float x0 = swap16_add(w0, w1), x1 = swap16_add(w2, w3), x2 = swap16_add(w4, 0.0f);
    const float y0 = pair<0x128>(x0, x1, hi8);    // row_ror 8: lanes i <-> i + 8
    const float y1 = pair<0x128>(x2, 0.0f, hi8);
    float z = pair<0x141>(y0, y1, hi4);           // row_half_mirror: lanes i <-> 7 - i
    z += dpp_f32<0xb1>(z);
    z += dpp_f32<0x4e>(z);
    return z;
  }

  // The value index (0..8) this lane stores after reduce(), or -1: one lane
  // per quad, the quads that hold a value.
  __device__ __forceinline__ int slot() const {
    const float z = reduce(1.0f, 2.0f, 3.0f, 4.0f, 5.0f, 6.0f, 7.0f, 8.0f, 9.0f);
    const int k = int(z * (1.0f / 64.0f) + 0.5f) - 1;
    return (threadIdx.x & 3u) == 0u ? k : -1;
  }
};

// Every lane takes the previous lane's v (DPP wave_shr:1); lane 0, which has
// no previous lane, takes `first`.
__device__ __forceinline__ float wave_shift_in(float first, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, first),
                                                               __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}

// A wave-uniform 64-bit value (read from LDS or computed the same in every
// lane) moved to scalar registers, so loops over it run on the scalar unit.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
  return (uint64_t(hi) << 32) | lo;
}

}  // namespace gsd
