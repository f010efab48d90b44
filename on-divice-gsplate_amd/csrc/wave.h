// Wave64 reductions for gfx950: DPP row rotations inside each 16-lane row,
// then the gfx950 permlane16/permlane32 swaps across rows.  Every lane ends
// with the full sum.
#pragma once

#include <hip/hip_runtime.h>

namespace gsd {

template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xb1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f32<0x4e>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f32<0x124>(v);  // row_ror 4
  v += dpp_f32<0x128>(v);  // row_ror 8: every lane holds its row's sum
  {
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, v), __builtin_bit_cast(unsigned, v),
                                                    false, false);
    v = __builtin_bit_cast(float, unsigned(r[0])) + __builtin_bit_cast(float, unsigned(r[1]));
  }
  {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, v), __builtin_bit_cast(unsigned, v),
                                                    false, false);
    v = __builtin_bit_cast(float, unsigned(r[0])) + __builtin_bit_cast(float, unsigned(r[1]));
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o, 64);
    v = t > v ? t : v;
  }
  return v;
}

}  // namespace gsd
