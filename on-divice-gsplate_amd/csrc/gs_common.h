// Device-side definitions shared by the gfx950 kernels: constants, the
// portable exp/log used on the bit-exact projection chain, the real SH basis,
// the opacity-aware footprint, and the raster record layout.
//
// Bit-exactness: the projection chain (camera transform, covariance, conic,
// footprint tile count, depth key) is evaluated with FP contraction off and
// only correctly rounded operations (+ - * /, sqrt, rint, floor, ceil) plus
// gs_expf / gs_logf below, which are built from those.  The CPU oracle
// restates the same sequence in C with contraction off, so both produce the
// same bits and the same tile sets (engine/shaders/train_common.glsl:163-213
// makes the same demand of its two passes with `precise`).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_hip.h"

namespace gsd {

// A pointer read from a device table (ViewDev, SceneDev, a sort segment)
// re-typed as global memory.  Accesses through the generic pointer compile
// to flat instructions, which count against both the vector-memory and the
// LDS/scalar wait counters, so the wait for the next table entry's scalar
// load, or for an LDS read, also waits for every flat load, store or atomic
// in flight; global instructions count only against the vector counter.
#define GS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GS_GLOBAL T* gmem(T* p) {
  return (GS_GLOBAL T*)p;
}
// Float add without a return value (the hardware atomic), through a global pointer.
__device__ __forceinline__ void global_add(GS_GLOBAL float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Vector loads and stores through the global view (HIP's vector types only
// bind generic references, so these go through the native vector types).
typedef float gs_f4v __attribute__((ext_vector_type(4)));
typedef unsigned int gs_u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 gload(const float4* p) {
  const gs_f4v v = *(const GS_GLOBAL gs_f4v*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
typedef unsigned int gs_u4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 gload(const uint4* p) {
  const gs_u4v v = *(const GS_GLOBAL gs_u4v*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gload(const uint2* p) {
  const gs_u2v v = *(const GS_GLOBAL gs_u2v*)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void gstore(float4* p, float4 v) {
  gs_f4v w;
  w.x = v.x, w.y = v.y, w.z = v.z, w.w = v.w;
  *(GS_GLOBAL gs_f4v*)p = w;
}

constexpr int kTile = GS_HIP_TILE;
constexpr float kNear = 0.02f;        // train_common.glsl:65
constexpr float kFar = 1000.0f;       // train_common.glsl:66
constexpr float kDilation = 0.3f;     // train_common.glsl:67
constexpr float kMinAlpha = 1.0f / 255.0f;
constexpr float kMaxAlpha = 0.99f;
constexpr float kMinTransmittance = 1e-4f;

// SSIM window: 11 taps of a Gaussian with sigma 1.5, normalised
// (train_common.glsl:59-62 lists the same values).
__constant__ static const float kSsimWindow[11] = {0.001028380f, 0.007598758f, 0.036000772f, 0.109360690f,
                                                   0.213005538f, 0.266011725f, 0.213005538f, 0.109360690f,
                                                   0.036000772f, 0.007598758f, 0.001028380f};
constexpr int kSsimRadius = 5;

__device__ __forceinline__ int sh_rest_coeffs(uint32_t degree) {
  return degree == 0u ? 0 : int((degree + 1u) * (degree + 1u)) - 1;
}

// ---- exp / log from correctly rounded operations --------------------------
// exp: Cody-Waite reduction x = n ln2 + r (ln2 split in a 16-bit head and a
// tail, so n * head is exact), e^r by its Taylor series to r^7 (|r| <= 0.35:
// truncation below 6e-9 relative), then 2^n applied as two exact power-of-two
// products.
__device__ __forceinline__ float gs_expf(float x) {
#pragma clang fp contract(off)
  if (x > 88.72f) return __int_as_float(0x7f800000);
  if (x < -103.97f) return 0.0f;
  const float n = rintf(x * 1.44269504f);
  float r = x - n * 0.693145751953125f;
  r = r - n * 1.428606765330187e-06f;
  float p = 1.98412698e-4f;  // 1/5040
  p = p * r + 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  const int k = int(n);
  const int k1 = k / 2, k2 = k - k1;
  p = p * __int_as_float((k1 + 127) << 23);
  return p * __int_as_float((k2 + 127) << 23);
}

// log of a positive normal float: x = m 2^e with m in [sqrt(1/2), sqrt(2)],
// log m = 2 atanh(s), s = (m - 1) / (m + 1), by the atanh series to s^11
// (|s| <= 0.172: truncation below 1e-10 relative).
__device__ __forceinline__ float gs_logf(float x) {
#pragma clang fp contract(off)
  const int bits = __float_as_int(x);
  int e = ((bits >> 23) & 0xff) - 127;
  float m = __int_as_float((bits & 0x007fffff) | 0x3f800000);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = 9.09090909e-2f;  // 1/11
  p = p * s2 + 1.11111111e-1f;
  p = p * s2 + 1.42857143e-1f;
  p = p * s2 + 2.0e-1f;
  p = p * s2 + 3.33333333e-1f;
  p = p * s2 + 1.0f;
  const float lm = (2.0f * s) * p;
  const float fe = float(e);
  return fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
}

__device__ __forceinline__ float gs_sigmoid(float x) {
#pragma clang fp contract(off)
  return 1.0f / (1.0f + gs_expf(-x));
}

// ---- real SH basis to degree 3 (3DGS sign convention) ---------------------
// Index 0 is DC; the constants are the standard real SH normalisations.
__device__ __forceinline__ void sh_basis(float x, float y, float z, uint32_t degree, float b[16]) {
#pragma clang fp contract(off)
  b[0] = 0.28209479177387814f;
  if (degree >= 1u) {
    b[1] = -0.4886025119029199f * y;
    b[2] = 0.4886025119029199f * z;
    b[3] = -0.4886025119029199f * x;
  }
  if (degree >= 2u) {
    const float xx = x * x, yy = y * y, zz = z * z;
    b[4] = 1.0925484305920792f * x * y;
    b[5] = -1.0925484305920792f * y * z;
    b[6] = 0.31539156525252005f * (2.0f * zz - xx - yy);
    b[7] = -1.0925484305920792f * x * z;
    b[8] = 0.5462742152960396f * (xx - yy);
    if (degree >= 3u) {
      b[9] = -0.5900435899266435f * y * (3.0f * xx - yy);
      b[10] = 2.890611442640554f * x * y * z;
      b[11] = -0.4570457994644658f * y * (4.0f * zz - xx - yy);
      b[12] = 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      b[13] = -0.4570457994644658f * x * (4.0f * zz - xx - yy);
      b[14] = 1.445305721320277f * z * (xx - yy);
      b[15] = -0.5900435899266435f * x * (xx - 3.0f * yy);
    }
  }
}

// d basis_k / d(direction) for k = 1..15 (gx, gy, gz per k).
__device__ __forceinline__ void sh_basis_grad(float x, float y, float z, uint32_t degree, float gx[16], float gy[16],
                                              float gz[16]) {
  const float C1 = 0.4886025119029199f;
  gx[1] = 0.0f, gy[1] = -C1, gz[1] = 0.0f;
  gx[2] = 0.0f, gy[2] = 0.0f, gz[2] = C1;
  gx[3] = -C1, gy[3] = 0.0f, gz[3] = 0.0f;
  if (degree >= 2u) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float A = 1.0925484305920792f, B = 0.31539156525252005f, C = 0.5462742152960396f;
    gx[4] = A * y, gy[4] = A * x, gz[4] = 0.0f;
    gx[5] = 0.0f, gy[5] = -A * z, gz[5] = -A * y;
    gx[6] = -2.0f * B * x, gy[6] = -2.0f * B * y, gz[6] = 4.0f * B * z;
    gx[7] = -A * z, gy[7] = 0.0f, gz[7] = -A * x;
    gx[8] = 2.0f * C * x, gy[8] = -2.0f * C * y, gz[8] = 0.0f;
    if (degree >= 3u) {
      const float D = -0.5900435899266435f, E = 2.890611442640554f, F = -0.4570457994644658f;
      const float G = 0.3731763325901154f, H = 1.445305721320277f;
      gx[9] = D * 6.0f * x * y, gy[9] = D * (3.0f * xx - 3.0f * yy), gz[9] = 0.0f;
      gx[10] = E * y * z, gy[10] = E * x * z, gz[10] = E * x * y;
      gx[11] = F * (-2.0f * x * y), gy[11] = F * (4.0f * zz - xx - 3.0f * yy), gz[11] = F * 8.0f * y * z;
      gx[12] = G * (-6.0f * x * z), gy[12] = G * (-6.0f * y * z), gz[12] = G * (6.0f * zz - 3.0f * xx - 3.0f * yy);
      gx[13] = F * (4.0f * zz - 3.0f * xx - yy), gy[13] = F * (-2.0f * x * y), gz[13] = F * 8.0f * x * z;
      gx[14] = H * 2.0f * x * z, gy[14] = H * (-2.0f * y * z), gz[14] = H * (xx - yy);
      gx[15] = D * (3.0f * xx - 3.0f * yy), gy[15] = D * (-6.0f * x * y), gz[15] = 0.0f;
    }
  }
}

// ---- opacity-aware footprint ---------------------------------------------
// A pixel centre p gets alpha >= 1/255 exactly where d^T cov^-1 d <= 2 ln(255 o),
// d = mean - p.  The threshold carries a hair of slack for rounding
// (train_common.glsl:163-172).
__device__ __forceinline__ float footprint_threshold(float opacity) {
#pragma clang fp contract(off)
  const float t = 2.0f * gs_logf(opacity * 255.0f);
  return t * 1.0001f + 1e-4f;
}

// [first, last] of the pixel indices in [0, size) whose centres x + 0.5 lie
// in [lo + 0.5, hi + 0.5]; first > last if none.  NaN clamps to the edge.
__device__ __forceinline__ void footprint_pixels(float lo, float hi, uint32_t size, int& first, int& last) {
  const int n = int(size);
  const float fn = float(n);
  first = min(max(int(fminf(fmaxf(ceilf(lo), -1.0f), fn)), 0), n);
  last = min(max(int(fminf(fmaxf(floorf(hi), -1.0f), fn)), -1), n - 1);
}

// Pixel rows the ellipse (centre m, dilated covariance a b c, threshold t) reaches.
__device__ __forceinline__ void footprint_rows(float my, float c, float t, uint32_t height, int& first, int& last) {
#pragma clang fp contract(off)
  const float ey = sqrtf(t * c);
  const float v = my - 0.5f;
  footprint_pixels(v - ey, v + ey, height, first, last);
}

// The terms of an ellipse's row spans that do not depend on the row: centre
// (u, v) in pixel-index coordinates, k = b/c, w = det/c, the height of the
// rightmost point b sqrt(t/a), and c, t.
struct RowSpan {
  float u, v, k, w, side, c, t;
};

__device__ __forceinline__ RowSpan row_span(float mx, float my, float a, float b, float c, float t) {
#pragma clang fp contract(off)
  RowSpan s;
  s.u = mx - 0.5f;
  s.v = my - 0.5f;
  s.k = b / c;
  s.w = (a * c - b * b) / c;
  s.side = b * sqrtf(t / a);
  s.c = c;
  s.t = t;
  return s;
}

// Bins [x0, x1) of bin row `ty` (bins of 2^side_log2 pixels: 16 for the
// raster tiles) that the ellipse reaches within pixel rows [row_first,
// row_last].  At height dy the ellipse spans
// dx = (b/c) dy -+ sqrt((det/c)(t - dy^2/c)); over [y0, y1] the right end is
// largest at the clamped height of the rightmost point, b sqrt(t/a), and the
// left end smallest at that of the leftmost, -b sqrt(t/a).
__device__ __forceinline__ void span_row_tiles(const RowSpan& s, int row_first, int row_last, uint32_t ty,
                                               uint32_t width, uint32_t& x0, uint32_t& x1, int side_log2 = 4) {
#pragma clang fp contract(off)
  const int top = int(ty) << side_log2;
  const float y0 = float(max(top, row_first)) - s.v;
  const float y1 = float(min(top + (1 << side_log2) - 1, row_last)) - s.v;
  const float dr = fminf(fmaxf(s.side, y0), y1);
  const float dl = fminf(fmaxf(-s.side, y0), y1);
  const float x_max = s.k * dr + sqrtf(fmaxf(s.w * (s.t - dr * dr / s.c), 0.0f));
  const float x_min = s.k * dl - sqrtf(fmaxf(s.w * (s.t - dl * dl / s.c), 0.0f));
  int first, last;
  footprint_pixels(s.u + x_min, s.u + x_max, width, first, last);
  if (first > last) {
    x0 = x1 = 0u;
    return;
  }
  x0 = uint32_t(first) >> side_log2;
  x1 = (uint32_t(last) >> side_log2) + 1u;
}

// ---- raster record --------------------------------------------------------
// 12 floats (48 B) per slot, read by the forward and backward rasterizers:
//   [0] mean x  [1] mean y  [2] conic a  [3] conic b
//   [4] conic c [5] opacity [6] r        [7] g
//   [8] b       [9] bbox lo (x | y << 16)  [10] bbox hi (x | y << 16)
//   [11] radius << 3 | clamped-channel bits
// The bbox is the pixel range of the footprint ellipse's bounding box.
struct __align__(16) Record {
  float4 a;  // mean xy, conic ab
  float4 b;  // conic c, opacity, rg
  float4 c;  // b, bbox lo, bbox hi, radius|clamp
};

// Pixel bbox test of a record against an inclusive pixel rectangle.
__device__ __forceinline__ bool bbox_hits(uint32_t lo, uint32_t hi, int rx0, int ry0, int rx1, int ry1) {
  const int bx0 = int(lo & 0xffffu), by0 = int(lo >> 16), bx1 = int(hi & 0xffffu), by1 = int(hi >> 16);
  return bx0 <= rx1 && bx1 >= rx0 && by0 <= ry1 && by1 >= ry0;
}

}  // namespace gsd
