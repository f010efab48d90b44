// Tile rasterizer: front-to-back compositing (forward) and the
// recompute-transmittance adjoint (backward), one 16x16 tile per workgroup of
// 256 threads.  Each wave owns an 8x8 quadrant of the tile.
//
// Both kernels stage the tile's splats through LDS in batches of 256 (one
// coalesced 48 B record load per thread).  The staging thread rewrites the
// record into evaluation form, (mean, A, B), (C, opacity, r, g), b with
// (A, B, C) = -(a/2, b, c/2), so the per-pixel exponent is
// power = dx (A dx + B dy) + C dy^2 (the same quadratic form as the
// reference's -(a dx^2 + c dy^2)/2 - b dx dy), and tests the footprint bbox
// against the four quadrants: a ballot per quadrant gives each wave a 64-bit
// mask of the batch's splats that can reach its pixels.  The mask is moved to
// scalar registers, so the walk over its bits is scalar control flow and the
// vector unit only evaluates splats.  (Scaling (A, B, C) by log2(e) to save
// exp's multiply moves the power by a few ulps of its terms, which for long
// thin splats is more than the oracle's rounding model of the power allows.)
//
// Forward (engine/shaders/train_forward.comp:59-141): alpha = min(0.99,
// o exp(power)), skipped below 1/255 or for power > 0; a pixel stops before
// the splat that would take T below 1e-4.  Writes IMAGE (rgb, final T) and
// LAST (1-based position of the last composited splat in the tile's range).
//
// Backward (the math of train_common.glsl:215-253 splat_pixel_grad, the loop
// of train_backward_serial.glsl): each pixel starts at its final T and walks
// its tile's list back to front from its last contributor, recovering
// T_i = T_{i+1} / (1 - alpha_i) and the colour behind each splat; nothing
// per pixel is stored by the forward besides T, LAST and dL/dC.  A pixel's
// share of the splat's mean and conic gradients is dL/dpower times a
// polynomial in (dx, dy) whose coefficients are the splat's, and
// dL/dpower = o dL/do, so each pixel adds only g dx, g dy, g dx^2, g dx dy,
// g dy^2 with g = dL/do; the splat's constants are applied once to the
// tile's sums.  The nine per-splat sums are summed over each wave by a transposed reduction
// (one quad ends with each value), over the four waves in LDS (ds_add_f32),
// and added to the slot's GRAD2D with one global float atomic each.  GRAD2D
// keeps a slot's nine sums together and the flush gives consecutive lanes
// consecutive (splat, value) entries, so one atomic instruction covers about
// seven slots' cache lines instead of 64.
//
// Work: the (view, tile) items of all views of the step in longest-list-first
// order, in 2x2 quads of neighbouring tiles.  Either one workgroup per item,
// or a persistent grid of a few workgroups per CU that take items from a
// device counter until none are left.  With one workgroup per item the
// hardware deals workgroups to the 8 XCDs in turn (workgroup b to XCD b mod
// 8); workgroup b takes item 32 (b / 32) + 4 (b mod 8) + (b / 8) mod 4, so
// the four tiles of a quad run on one XCD at about the same time and share
// their records through its L2, while the list is still taken in order
// 32 items at a time.
#include <algorithm>

#include "cull.h"
#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;

constexpr int kThreads = 256;
constexpr int kBatch = 256;
constexpr int kGrad = GS_HIP_GRAD2D_FLOATS;
constexpr uint32_t kBucketWavesPerCu = 20;  // stored-state backward: one-wave workgroups per CU (LDS allows 21)

struct RasterArgs {
  const gsh::ViewDev* __restrict__ views;
  const uint32_t* __restrict__ work;  // view << 16 | tile, longest lists first
  uint32_t* __restrict__ counter;     // persistent mode: next work item
  uint32_t items;                     // every view's tiles
  uint32_t persistent;
  // Stored-state backward: the forward lists every walked bucket here.
  gsh::BucketEntry* __restrict__ bucket_list;
  uint32_t* __restrict__ bucket_counts;  // [0] listed by the forward, [1] taken by the backward
  uint64_t bucket_entries;
};

struct PixelCoord {
  int px, py;  // pixel
  int tx0, ty0;  // tile origin
  bool inside;
  uint32_t tile;
  int wave, lane;  // wave is scalar
};

__device__ __forceinline__ PixelCoord pixel_coord(const gs_hip_frame& f, uint32_t tile) {
  PixelCoord c;
  c.tile = tile;
  c.tx0 = int(tile % f.tiles_x) * kTile;
  c.ty0 = int(tile / f.tiles_x) * kTile;
  c.wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  c.lane = int(threadIdx.x & 63u);
  c.px = c.tx0 + (c.wave & 1) * 8 + (c.lane & 7);
  c.py = c.ty0 + (c.wave >> 1) * 8 + (c.lane >> 3);
  c.inside = c.px < int(f.cam.width) && c.py < int(f.cam.height);
  return c;
}

// One batch of splats in LDS, in evaluation form: three float4 per splat,
// so one address register and three immediate offsets reach a splat.
template <int B = kBatch>
struct Staging {
  // [3j]: mean x, mean y, A, B; [3j + 1]: C, opacity, r, g; [3j + 2].x: b;
  // (A, B, C) = -(a/2, b, c/2)
  float4 sp[B * 3];
  uint64_t mask[4][B / 64];  // [quadrant][loading wave]
  __device__ __forceinline__ float4 a(int j) const { return sp[3 * j]; }
  __device__ __forceinline__ float4 b(int j) const { return sp[3 * j + 1]; }
  __device__ __forceinline__ float blue(int j) const { return sp[3 * j + 2].x; }
};

// Loads pair k's record into slot t and returns which quadrants it can draw
// in (bit q): its footprint bbox meets the quadrant and, at some point of the
// quadrant's pixel-centre rectangle, o exp(-q) reaches 1/255 (q the conic's
// quadratic form), with a margin above the rounding of the per-pixel
// evaluation (gsh::quadrant_may_draw), so a quadrant is dropped only where
// every pixel skips the splat anyway.
template <int B>
__device__ __forceinline__ uint32_t stage(Staging<B>& s, const float4* __restrict__ records, uint32_t g, int t, int tx0,
                                          int ty0) {
  const float4* r = records + size_t(g) * 3u;
  const float4 ra = gload(r), rb = gload(r + 1), rc = gload(r + 2);
  s.sp[3 * t] = make_float4(ra.x, ra.y, -0.5f * ra.z, -ra.w);
  s.sp[3 * t + 1] = make_float4(-0.5f * rb.x, rb.y, rb.z, rb.w);
  s.sp[3 * t + 2].x = rc.x;
  const uint32_t lo = __float_as_uint(rc.y), hi = __float_as_uint(rc.z);
  uint32_t hits = 0u;
  for (int q = 0; q < 4; ++q) {
    const int qx = tx0 + (q & 1) * 8, qy = ty0 + (q >> 1) * 8;
    hits |= bbox_hits(lo, hi, qx, qy, qx + 7, qy + 7) ? (1u << q) : 0u;
  }
  if (hits != 0u) {
    const float a = ra.z, b = ra.w, c = rb.x;
    const float limit = __logf(255.0f * rb.y);  // drawn where q <= ln(255 o)
    // The clamped vertices need only be near the edges' minima: the hardware
    // reciprocal will do.
    const float ia = __builtin_amdgcn_rcpf(a), ic = __builtin_amdgcn_rcpf(c);
    for (int q = 0; q < 4; ++q) {
      if (((hits >> q) & 1u) == 0u) continue;
      const float x0 = float(tx0 + (q & 1) * 8) + 0.5f, y0 = float(ty0 + (q >> 1) * 8) + 0.5f;
      if (!gsh::quadrant_may_draw(a, b, c, ia, ic, limit, ra.x - (x0 + 7.0f), ra.x - x0, ra.y - (y0 + 7.0f),
                                  ra.y - y0))
        hits &= ~(1u << q);
    }
  }
  return hits;
}

template <int B>
__device__ __forceinline__ void publish_masks(Staging<B>& s, uint32_t hits, int wave, int lane) {
  for (int q = 0; q < 4; ++q) {
    const uint64_t m = __ballot((hits >> q) & 1u);
    if (lane == 0) s.mask[q][wave] = m;
  }
}

// Power and alpha of a splat at a pixel centre, shared by both kernels so
// they make the same skip decisions on the same bits.
struct SplatEval {
  float dx, dy, power, G, raw, alpha;  // G = exp(power)
};

__device__ __forceinline__ SplatEval eval_splat(float4 sa, float4 sb, float cx, float cy) {
  SplatEval e;
  e.dx = sa.x - cx;
  e.dy = sa.y - cy;
  const float u = fmaf(sa.w, e.dy, sa.z * e.dx);  // A dx + B dy
  e.power = fmaf(e.dx, u, (sb.x * e.dy) * e.dy);
  e.G = __expf(e.power);
  e.raw = sb.y * e.G;
  e.alpha = fminf(kMaxAlpha, e.raw);
  return e;
}

// Front-to-back compositing of one tile of one view, staging B splats per
// batch (B / 256 per thread).  kStore (the stored-state backward): at list
// position 64 k, k >= 1, each wave that still composites stores its pixels'
// (C, T) in the tile's bucket slot k, and the tile's buckets (as far as its
// pixels' last contributors reach) are listed for the backward.
template <int B, bool kStore>
__device__ __forceinline__ void forward_tile(const gsh::ViewDev& a, uint32_t item, Staging<B>& s,
                                             const RasterArgs& ra, uint32_t* s_last) {
  const gs_hip_frame& f = a.f;
  const uint32_t tile = item & gsh::kMaxTilesPerView;
  const PixelCoord pc = pixel_coord(f, tile);
  const uint32_t bin = gsh::bin_of_tile(a, pc.tile);
  const uint2 range = gload(a.ranges + bin);
  const float cx = float(pc.px) + 0.5f, cy = float(pc.py) + 0.5f;
  // Ts is T while the pixel composites and -T once it has stopped (or lies
  // outside the image): "not done" is Ts > 0, one compare whose lane mask is
  // the wave's ballot.
  float Ts = pc.inside ? 1.0f : -1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  uint32_t last = 0u;

  for (uint32_t start = range.x; start < range.y; start += B) {
    if (__syncthreads_and(Ts < 0.0f)) break;
    const uint32_t n = min(uint32_t(B), range.y - start);
    for (int r = 0; r < B / kThreads; ++r) {
      const uint32_t t = uint32_t(r * kThreads) + threadIdx.x;
      uint32_t hits = 0u;
      if (t < n) hits = stage(s, a.records, gmem(a.pair_gaussians)[start + t], int(t), pc.tx0, pc.ty0);
      publish_masks(s, hits, r * (kThreads / 64) + pc.wave, pc.lane);
    }
    __syncthreads();
    // The walk ends with the masks or when every pixel of the wave is done;
    // both loop conditions are wave-uniform (scalar registers).
    uint64_t alive = __ballot(Ts > 0.0f);
    const uint32_t base = start - range.x + 1u;  // LAST of splat j is base + j
    for (int chunk = 0; chunk < B / 64 && alive != 0ull; ++chunk) {
      if constexpr (kStore) {
        const uint32_t k = (start - range.x) / gsh::kBucketSplats + uint32_t(chunk);
        if (k != 0u && uint32_t(chunk) * 64u < n) {
          const uint64_t slot = gsh::bucket_slot(range.x, bin, k, tile, f.tiles_x, a.bin_shift);
          gstore(a.bucket_state + slot * gsh::kBucketPixels + uint32_t(pc.wave * 64 + pc.lane),
                 make_float4(cr, cg, cb, Ts));
        }
      }
      uint64_t m = uniform_u64(s.mask[pc.wave][chunk]);
      while (m != 0ull && alive != 0ull) {
        const int j = chunk * 64 + __builtin_ctzll(m);
        m &= m - 1ull;
        // Branch-free: lanes that are done or skip this splat add zero.
        const float4 sb = s.b(j);
        const float blue = s.blue(j);
        const SplatEval e = eval_splat(s.a(j), sb, cx, cy);
        const float next_T = Ts * (1.0f - e.alpha);
        const bool ok = Ts > 0.0f && e.power <= 0.0f && e.alpha >= kMinAlpha;
        const bool stop = ok && next_T < kMinTransmittance;
        const bool take = ok && !stop;
        const float w = take ? e.alpha * Ts : 0.0f;
        cr = fmaf(sb.z, w, cr);
        cg = fmaf(sb.w, w, cg);
        cb = fmaf(blue, w, cb);
        last = take ? base + uint32_t(j) : last;
        Ts = take ? next_T : (stop ? -Ts : Ts);
        alive = __ballot(Ts > 0.0f);
      }
    }
    __syncthreads();
  }
  const float T = fabsf(Ts);
  if (pc.inside) {
    const uint32_t pix = uint32_t(pc.py) * f.cam.width + uint32_t(pc.px);
    gstore(a.image + pix, make_float4(cr, cg, cb, T));
    gmem(a.last)[pix] = last;
  }
  if constexpr (kStore) {
    const uint32_t wl = wave_max_u32(last);
    if (pc.lane == 0) s_last[pc.wave] = wl;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tile_last = max(max(s_last[0], s_last[1]), max(s_last[2], s_last[3]));
      const uint32_t buckets = (tile_last + gsh::kBucketSplats - 1u) / gsh::kBucketSplats;
      if (buckets != 0u) {
        const uint32_t first = atomicAdd(ra.bucket_counts, buckets);
        for (uint32_t k = 0; k < buckets; ++k) {
          if (uint64_t(first) + k < ra.bucket_entries) ra.bucket_list[first + k] = gsh::BucketEntry{item, k};
        }
      }
    }
  }
}

// kDet: the deterministic variant keeps each wave's sums apart ([wave][splat][9])
// and adds them in wave order, then adds the tile's sums to the slot as 64-bit
// fixed point (integer additions commute, so the result does not depend on
// which tile or wave came first).
template <bool kDet>
struct BackwardShared {
  Staging<> s;
  uint32_t g[kBatch];
  float grad[(kDet ? 4 : 1) * kBatch * kGrad];  // [wave][splat][9]: the nine lanes of one reduction hit nine banks
  uint32_t wave_last[4];
};

// Fixed point of the deterministic sums: 2^-44 resolution, |value| < 2^18.
constexpr float kFixedScale = 17592186044416.0f;  // 2^44
constexpr double kFixedUnit = 1.0 / 17592186044416.0;
constexpr float kFixedLimit = 262144.0f;  // 2^18

__device__ __forceinline__ long long to_fixed(float v) {
  return __float2ll_rn(fminf(fmaxf(v, -kFixedLimit), kFixedLimit) * kFixedScale);
}

// The tile's sum of the per-pixel value c of staged splat j (the waves'
// sums added in wave order for the deterministic variant).
template <bool kDet>
__device__ __forceinline__ float tile_sum(const float* s_grad, uint32_t j, uint32_t c) {
  const uint32_t e = j * uint32_t(kGrad) + c;
  if (!kDet) return s_grad[e];
  constexpr uint32_t w = kBatch * kGrad;
  return (s_grad[e] + s_grad[w + e]) + (s_grad[2 * w + e] + s_grad[3 * w + e]);
}

// GRAD2D value c of staged splat j from the tile's sums of g dx, g dy,
// g dx^2, g dx dy, g dy^2, g (g = dL/do) and the colour gradient: with
// power = A dx^2 + B dx dy + C dy^2, (A, B, C) = -(a/2, b, c/2), and
// dL/dpower = o g,
//   dL/dmean = o (sum g d power/d mean) = o (2A sx + B sy, B sx + 2C sy),
//   dL/d(a, b, c) = -o (sxx / 2, sxy, syy / 2).
template <bool kDet>
__device__ __forceinline__ float tile_gradient(const Staging<>& s, const float* s_grad, uint32_t j, uint32_t c) {
  const float sum = tile_sum<kDet>(s_grad, j, c);
  if (c >= 5u) return sum;
  const float4 sa = s.a(int(j));
  const float4 sb = s.b(int(j));
  const float o = sb.y;
  if (c == 0u) return o * fmaf(2.0f * sa.z, sum, sa.w * tile_sum<kDet>(s_grad, j, 1u));
  if (c == 1u) return o * fmaf(2.0f * sb.x, sum, sa.w * tile_sum<kDet>(s_grad, j, 0u));
  return (c == 3u ? -o : -0.5f * o) * sum;
}

// Recompute-transmittance adjoint of one tile of one view.
template <bool kDet>
__device__ __forceinline__ void backward_tile(const gsh::ViewDev& a, uint32_t tile, BackwardShared<kDet>& sh,
                                              const Reduce9& red, int slot) {
  const gs_hip_frame& f = a.f;
  Staging<>& s = sh.s;
  uint32_t* s_g = sh.g;
  float* s_grad = sh.grad;
  uint32_t* s_wave_last = sh.wave_last;
  const PixelCoord pc = pixel_coord(f, tile);
  const uint2 range = gload(a.ranges + gsh::bin_of_tile(a, pc.tile));
  const float cx = float(pc.px) + 0.5f, cy = float(pc.py) + 0.5f;

  float T_final = 0.0f, gr = 0.0f, gg = 0.0f, gb = 0.0f;
  uint32_t last = 0u;
  if (pc.inside) {
    const uint32_t pix = uint32_t(pc.py) * f.cam.width + uint32_t(pc.px);
    T_final = gload(a.image + pix).w;
    last = gmem(a.last)[pix];
    const float4 d = gload(a.dldc + pix);
    gr = d.x, gg = d.y, gb = d.z;
  }
  // T_final/(1 - alpha) times bg . dL/dC: the background seen through the splat.
  const float bg_term = T_final * (f.background[0] * gr + f.background[1] * gg + f.background[2] * gb);
  // Back-to-front state: T behind the current splat, and the colour behind
  // it carried as its dot product with dL/dC (dL/dC is constant along the
  // walk, so acc' = a c + (1 - a) acc gives acc'.g = a (c.g) + (1 - a) acc.g:
  // one running value instead of three channels).
  float T = T_final, acc_dot = 0.0f;

  const uint32_t wl = wave_max_u32(last);
  if (pc.lane == 0) s_wave_last[pc.wave] = wl;
  __syncthreads();
  const uint32_t tile_last = max(max(s_wave_last[0], s_wave_last[1]), max(s_wave_last[2], s_wave_last[3]));
  const uint32_t end = range.x + tile_last;

  for (uint32_t bend = end; bend > range.x;) {
    const uint32_t bstart = bend > range.x + kBatch ? bend - kBatch : range.x;
    const uint32_t n = bend - bstart;
    uint32_t hits = 0u;
    if (threadIdx.x < n) {
      const uint32_t k = bstart + threadIdx.x;
      const uint32_t g = gmem(a.pair_gaussians)[k];
      s_g[threadIdx.x] = g;
      for (int w = 0; w < (kDet ? 4 : 1); ++w)
        for (int c = 0; c < kGrad; ++c) s_grad[(w * kBatch + threadIdx.x) * kGrad + c] = 0.0f;
      hits = stage(s, a.records, g, threadIdx.x, pc.tx0, pc.ty0);
      const uint32_t pos = k - range.x;
      for (int q = 0; q < 4; ++q)
        if (pos >= s_wave_last[q]) hits &= ~(1u << q);
    }
    publish_masks(s, hits, pc.wave, pc.lane);
    __syncthreads();
    const uint32_t base = bstart - range.x;  // position of splat j is base + j
    for (int chunk = kBatch / 64 - 1; chunk >= 0; --chunk) {
      uint64_t m = uniform_u64(s.mask[pc.wave][chunk]);
      while (m != 0ull) {
        const int bit = 63 - __builtin_clzll(m);
        m ^= 1ull << bit;
        const int j = chunk * 64 + bit;
        const float4 sa = s.a(j);
        const float4 sb = s.b(j);
        const float blue = s.blue(j);
        const SplatEval e = eval_splat(sa, sb, cx, cy);
        const float alpha = e.alpha;
        const bool contributes = base + uint32_t(j) < last && e.power <= 0.0f && alpha >= kMinAlpha;
        if (__ballot(contributes) == 0ull) continue;
        // Branch-free from here: lanes that do not contribute keep their
        // state and add zeros: their alpha counts as 0 where it scales, and
        // selects keep exp's possible inf out of the rest.
        const float a_eff = contributes ? alpha : 0.0f;
        const float inv = __builtin_amdgcn_rcpf(1.0f - alpha);
        const float T_in = T * inv;  // transmittance in front of this splat
        const float col_dot = fmaf(sb.z, gr, fmaf(sb.w, gg, blue * gb));  // colour . dL/dC
        const float behind = col_dot - acc_dot;
        const float dL_dalpha = behind * T_in - bg_term * inv;
        const bool unclamped = contributes && e.raw < kMaxAlpha;
        const float wT = a_eff * T_in;
        // dL/do = exp(power) dL/dalpha (alpha unclamped), and dL/dpower = o dL/do.
        const float g5 = unclamped ? e.G * dL_dalpha : 0.0f;
        const float gx = g5 * e.dx, gy = g5 * e.dy;
        const float g6 = wT * gr, g7 = wT * gg, g8 = wT * gb;
        T = contributes ? T_in : T;
        acc_dot = fmaf(a_eff, behind, acc_dot);  // this splat joins the colour behind the next
        const float z = red.reduce(gx, gy, gx * e.dx, gx * e.dy, gy * e.dy, g5, g6, g7, g8);
        if (slot >= 0) {
          if (kDet) {
            s_grad[(pc.wave * kBatch + j) * kGrad + slot] = z;  // this wave's only write of the value
          } else {
            __hip_atomic_fetch_add((__attribute__((address_space(3))) float*)s_grad + (uint32_t(j) * kGrad + uint32_t(slot)),
                                   z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < n * uint32_t(kGrad); e += kThreads) {
      const uint32_t j = e / uint32_t(kGrad), c = e - j * uint32_t(kGrad);
      const float v = tile_gradient<kDet>(s, s_grad, j, c);
      if (v == 0.0f) continue;
      if (kDet) {
        __hip_atomic_fetch_add(gmem(a.grad_fx) + size_t(s_g[j]) * kGrad + c, static_cast<unsigned long long>(to_fixed(v)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        global_add(gmem(a.grad2d) + size_t(s_g[j]) * kGrad + c, v);
      }
    }
    __syncthreads();
    bend = bstart;
  }
}

// Both kernels: one workgroup per work item, or (persistent) a grid that
// takes items from the counter until they run out; the exit test reads the
// shared item, so the whole workgroup leaves together.
__device__ __forceinline__ uint32_t next_item(const RasterArgs& a, uint32_t* s_item) {
  if (threadIdx.x == 0) *s_item = atomicAdd(a.counter, 1u);
  __syncthreads();
  const uint32_t item = __builtin_amdgcn_readfirstlane(*s_item);  // uniform: scalar loads from here on
  __syncthreads();
  return item;
}

// Workgroup b's item when each takes one (see the work order above).
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t items) {
  constexpr uint32_t kXcds = 8;
  constexpr uint32_t kRound = kXcds * 4u;
  if (b >= items - items % kRound) return b;
  return kRound * (b / kRound) + 4u * (b % kXcds) + (b / kXcds) % 4u;
}

template <int B, bool kStore>
__global__ void __launch_bounds__(kThreads) k_forward(RasterArgs a) {
  __shared__ Staging<B> s;
  __shared__ uint32_t s_item;
  __shared__ uint32_t s_last[4];
  uint32_t item = a.persistent != 0u ? next_item(a, &s_item) : xcd_item(blockIdx.x, a.items);
  while (item < a.items) {
    const uint32_t w = a.work[item];
    if (w != gsh::kNoWork) forward_tile<B, kStore>(a.views[w >> gsh::kTileBits], w, s, a, s_last);
    if (a.persistent == 0u) break;
    __syncthreads();
    item = next_item(a, &s_item);
  }
}

template <bool kDet>
__global__ void __launch_bounds__(kThreads) k_backward(RasterArgs a) {
  __shared__ BackwardShared<kDet> sh;
  __shared__ uint32_t s_item;
  const Reduce9 red;
  const int slot = red.slot();
  uint32_t item = a.persistent != 0u ? next_item(a, &s_item) : xcd_item(blockIdx.x, a.items);
  while (item < a.items) {
    const uint32_t w = a.work[item];
    if (w != gsh::kNoWork) backward_tile<kDet>(a.views[w >> gsh::kTileBits], w & gsh::kMaxTilesPerView, sh, red, slot);
    if (a.persistent == 0u) break;
    __syncthreads();
    item = next_item(a, &s_item);
  }
}

// Stored-state backward (train_backward_bucket.glsl's scheme, "Taming 3DGS",
// Mallick et al. 2024, laid out for wave64): one wave per bucket, i.e. 64
// consecutive splats of a tile's walk, lane l owning splat 64 k + l.  The
// tile's pixels that still composite at the bucket's start (LAST > 64 k) in
// quadrants some splat of the bucket can draw in pass through the lanes as a
// pipeline, front to back: at step s lane l works on pixel s - l and hands
// its state to lane l + 1 (DPP wave_shr), so each lane sums its own splat's
// gradient in registers and nothing is reduced across lanes.  A pixel enters
// with the state the forward stored at the bucket's start, T (exact: the
// forward's own recurrence) and R = dL/dC . (colour of everything from here
// on, background included) = dL/dC . (C_out - C_front); a splat it draws
// gives
//   dL/dalpha = T (dL/dC . c) - R' / (1 - alpha),  R' = R - (dL/dC . c) alpha T,
// and hands on T (1 - alpha) and R'.  The waves share nothing, so a
// workgroup is one wave; a persistent grid of them takes buckets from the
// list the forward wrote until it runs out.
struct BucketShared {
  union {
    float4 glast[gsh::kBucketPixels];        // pixel e of the feed: dL/dC, LAST (bits)
    float grad[gsh::kBucketSplats * kGrad];  // after the sweep: each splat's nine sums
  };
  float2 state[gsh::kBucketPixels];   // T, R at the bucket's start
  float2 centre[gsh::kBucketPixels];  // pixel centre
  uint32_t g[gsh::kBucketSplats];
};

__device__ __forceinline__ void backward_bucket(const gsh::ViewDev& a, uint32_t tile, uint32_t k, BucketShared& sh) {
  const gs_hip_frame& f = a.f;
  const int lane = int(threadIdx.x);
  const int tx0 = int(tile % f.tiles_x) * kTile, ty0 = int(tile / f.tiles_x) * kTile;
  const uint32_t bin = gsh::bin_of_tile(a, tile);
  const uint2 range = gload(a.ranges + bin);
  const uint32_t first = k * gsh::kBucketSplats;

  // The tile's pixels, quadrant q in register q (lane = its offset in the quadrant).
  uint32_t last[4], pix[4];
  bool inside[4];
  uint32_t most = 0u;
  for (int q = 0; q < 4; ++q) {
    const int px = tx0 + (q & 1) * 8 + (lane & 7), py = ty0 + (q >> 1) * 8 + (lane >> 3);
    inside[q] = px < int(f.cam.width) && py < int(f.cam.height);
    pix[q] = inside[q] ? uint32_t(py) * f.cam.width + uint32_t(px) : 0u;
    last[q] = inside[q] ? gmem(a.last)[pix[q]] : 0u;
    most = max(most, last[q]);
  }
  const uint32_t tile_last = __builtin_amdgcn_readfirstlane(wave_max_u32(most));

  // This lane's splat, in evaluation form (as staged by the tile kernels).
  const uint32_t pos = first + uint32_t(lane);
  const bool valid = pos < tile_last;
  float4 sa = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sb = sa;
  float blue = 0.0f;
  uint32_t g = 0u, hits = 0u;
  if (valid) {
    g = gmem(a.pair_gaussians)[range.x + pos];
    const float4* r = a.records + size_t(g) * 3u;
    const float4 ra = gload(r), rb = gload(r + 1), rc = gload(r + 2);
    sa = make_float4(ra.x, ra.y, -0.5f * ra.z, -ra.w);
    sb = make_float4(-0.5f * rb.x, rb.y, rb.z, rb.w);
    blue = rc.x;
    const uint32_t lo = __float_as_uint(rc.y), hi = __float_as_uint(rc.z);
    for (int q = 0; q < 4; ++q) {
      const int qx = tx0 + (q & 1) * 8, qy = ty0 + (q >> 1) * 8;
      hits |= bbox_hits(lo, hi, qx, qy, qx + 7, qy + 7) ? (1u << q) : 0u;
    }
    if (hits != 0u) {
      const float limit = __logf(255.0f * rb.y);
      const float ia = __builtin_amdgcn_rcpf(ra.z), ic = __builtin_amdgcn_rcpf(rb.x);
      for (int q = 0; q < 4; ++q) {
        if (((hits >> q) & 1u) == 0u) continue;
        const float x0 = float(tx0 + (q & 1) * 8) + 0.5f, y0 = float(ty0 + (q >> 1) * 8) + 0.5f;
        if (!gsh::quadrant_may_draw(ra.z, ra.w, rb.x, ia, ic, limit, ra.x - (x0 + 7.0f), ra.x - x0,
                                    ra.y - (y0 + 7.0f), ra.y - y0))
          hits &= ~(1u << q);
      }
    }
  }
  sh.g[lane] = g;

  // The feed: pixels that still composite at the bucket's start, in quadrants
  // some splat of the bucket may draw in.
  const float bg0 = f.background[0], bg1 = f.background[1], bg2 = f.background[2];
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t n = 0u;
  for (int q = 0; q < 4; ++q) {
    const bool drawn = __ballot((hits >> q) & 1u) != 0ull;
    const bool active = drawn && last[q] > first;
    const uint64_t m = __ballot(active);
    if (active) {
      const uint32_t e = n + uint32_t(__popcll(m & below));
      const float4 img = gload(a.image + pix[q]);
      const float4 d = gload(a.dldc + pix[q]);
      float4 front = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
      if (k != 0u) {
        const uint64_t slot = gsh::bucket_slot(range.x, bin, k, tile, f.tiles_x, a.bin_shift);
        front = gload(a.bucket_state + slot * gsh::kBucketPixels + uint32_t(q * 64 + lane));
      }
      const float out_dot = d.x * fmaf(img.w, bg0, img.x) + d.y * fmaf(img.w, bg1, img.y) + d.z * fmaf(img.w, bg2, img.z);
      const float front_dot = d.x * front.x + d.y * front.y + d.z * front.z;
      sh.glast[e] = make_float4(d.x, d.y, d.z, __uint_as_float(last[q]));
      sh.state[e] = make_float2(front.w, out_dot - front_dot);
      sh.centre[e] = make_float2(float(tx0 + (q & 1) * 8 + (lane & 7)) + 0.5f,
                                 float(ty0 + (q >> 1) * 8 + (lane >> 3)) + 0.5f);
    }
    n += uint32_t(__popcll(m));
  }
  __syncthreads();

  // The pipeline: at step s this lane holds pixel s - lane's T and R (shifted
  // in from the previous lane) and reads the pixel's constants from the feed.
  float T = 0.0f, R = 0.0f;
  float S0 = 0.0f, S1 = 0.0f, S2 = 0.0f, S3 = 0.0f, S4 = 0.0f, S5 = 0.0f, S6 = 0.0f, S7 = 0.0f, S8 = 0.0f;
  const uint32_t steps = n == 0u ? 0u : n + gsh::kBucketSplats - 1u;
  for (uint32_t s = 0; s < steps; ++s) {
    const float2 in = s < n ? sh.state[s] : make_float2(0.0f, 0.0f);
    T = wave_shift_in(in.x, T);
    R = wave_shift_in(in.y, R);
    const uint32_t e = s - uint32_t(lane);  // wraps for lanes not reached yet
    const bool here = e < n;
    const uint32_t ec = here ? e : 0u;
    const float4 gl = sh.glast[ec];
    const float2 c = sh.centre[ec];
    const SplatEval ev = eval_splat(sa, sb, c.x, c.y);
    // Non-short-circuit: the feed entry is read whether or not it is this lane's.
    const bool contributes = valid & here & (pos < __float_as_uint(gl.w)) & (ev.power <= 0.0f) &
                             (ev.alpha >= kMinAlpha);
    // Lanes that do not contribute add exact zeros (selects, not products: a
    // lane between pixels may hold anything).
    const float gc = fmaf(sb.z, gl.x, fmaf(sb.w, gl.y, blue * gl.z));  // colour . dL/dC
    const float w = contributes ? ev.alpha * T : 0.0f;
    const float rest = fmaf(-gc, w, R);
    const float inv = __builtin_amdgcn_rcpf(1.0f - ev.alpha);
    const float dL_dalpha = fmaf(T, gc, -rest * inv);
    const float g5 = contributes && ev.raw < kMaxAlpha ? ev.G * dL_dalpha : 0.0f;
    const float gx = g5 * ev.dx, gy = g5 * ev.dy;
    S0 += gx;
    S1 += gy;
    S2 = fmaf(gx, ev.dx, S2);
    S3 = fmaf(gx, ev.dy, S3);
    S4 = fmaf(gy, ev.dy, S4);
    S5 += g5;
    S6 = fmaf(w, gl.x, S6);
    S7 = fmaf(w, gl.y, S7);
    S8 = fmaf(w, gl.z, S8);
    T = contributes ? T * (1.0f - ev.alpha) : T;
    R = contributes ? rest : R;
  }

  // The splat's constants applied once (as tile_gradient), then the flush:
  // consecutive lanes add consecutive (splat, value) entries.
  __syncthreads();  // the feed is read; its memory takes the sums
  const float o = sb.y;
  float* out = sh.grad + lane * kGrad;
  out[0] = o * fmaf(2.0f * sa.z, S0, sa.w * S1);
  out[1] = o * fmaf(2.0f * sb.x, S1, sa.w * S0);
  out[2] = -0.5f * o * S2;
  out[3] = -o * S3;
  out[4] = -0.5f * o * S4;
  out[5] = S5;
  out[6] = S6;
  out[7] = S7;
  out[8] = S8;
  __syncthreads();
  const uint32_t splats = min(gsh::kBucketSplats, tile_last > first ? tile_last - first : 0u);
  for (uint32_t e = uint32_t(lane); e < splats * uint32_t(kGrad); e += 64u) {
    const uint32_t j = e / uint32_t(kGrad), c = e - j * uint32_t(kGrad);
    const float v = sh.grad[e];
    if (v != 0.0f) global_add(gmem(a.grad2d) + size_t(sh.g[j]) * kGrad + c, v);
  }
  __syncthreads();
}

// The next list entry for this wave (one atomic by lane 0, in a scalar register).
__device__ __forceinline__ uint32_t take_bucket(uint32_t* counter) {
  uint32_t v = 0u;
  if (threadIdx.x == 0) v = atomicAdd(counter, 1u);
  return __builtin_amdgcn_readfirstlane(v);
}

__global__ void __launch_bounds__(64) k_backward_stored(RasterArgs a) {
  __shared__ BucketShared sh;
  const uint32_t listed = uint32_t(min(uint64_t(gmem(a.bucket_counts)[0]), a.bucket_entries));
  for (;;) {
    const uint32_t taken = take_bucket(a.bucket_counts + 1);
    if (taken >= listed) break;
    const gsh::BucketEntry b = a.bucket_list[taken];
    const uint32_t item = __builtin_amdgcn_readfirstlane(b.item), k = __builtin_amdgcn_readfirstlane(b.k);
    backward_bucket(a.views[item >> gsh::kTileBits], item & gsh::kMaxTilesPerView, k, sh);
  }
}

// Deterministic backward: the fixed-point sums become GRAD2D and are cleared
// for the next step (they are zero between steps: cleared when made, and
// after every use).
__global__ void __launch_bounds__(256) k_grad_from_fixed(const gsh::ViewDev* __restrict__ views) {
  const gsh::ViewDev& v = views[blockIdx.y];
  const uint64_t n = uint64_t(v.f.slots) * kGrad;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    gmem(v.grad2d)[i] = float(double(static_cast<long long>(gmem(v.grad_fx)[i])) * kFixedUnit);
    gmem(v.grad_fx)[i] = 0ull;
  }
}

RasterArgs make_args(gs_hip_workspace* ws, uint32_t count, int which, uint32_t& grid) {
  RasterArgs a;
  a.views = ws->dev_views;
  a.work = ws->work;
  a.counter = ws->work_counters + which;
  a.items = ws->shape.work_items;
  a.bucket_list = ws->bucket_list;
  a.bucket_counts = ws->work_counters + 2;
  a.bucket_entries = ws->shape.stored_state ? ws->bucket_list_entries : 0;
  (void)count;
  a.persistent = ws->tuning.raster_groups_per_cu > 0 ? 1u : 0u;
  grid = a.persistent ? std::min<uint32_t>(a.items, ws->cus * ws->tuning.raster_groups_per_cu) : a.items;
  return a;
}

}  // namespace

namespace gsh {

int launch_forward(gs_hip_workspace* ws, uint32_t count) {
  uint32_t grid = 0;
  const RasterArgs a = make_args(ws, count, 0, grid);
  if (grid == 0) return 0;
  if (a.persistent != 0u) {
    const int rc = fill_u32(ws, a.counter, 1, 0u);
    if (rc != 0) return rc;
  }
  timer_begin(ws, "forward");
  if (ws->shape.stored_state) {
    k_forward<kBatch, true><<<grid, kThreads, 0, ws->stream>>>(a);
  } else {
    k_forward<kBatch, false><<<grid, kThreads, 0, ws->stream>>>(a);
  }
  GS_CHECK_LAUNCH("k_forward");
  timer_end(ws);
  return 0;
}

int launch_backward(gs_hip_workspace* ws, uint32_t count) {
  uint32_t grid = 0;
  const RasterArgs a = make_args(ws, count, 1, grid);
  if (grid == 0) return 0;
  if (ws->shape.stored_state) {
    // Waves take buckets from the list the forward wrote (its count was
    // cleared with the work order) until it runs out.
    const int rc = fill_u32(ws, a.bucket_counts + 1, 1, 0u);
    if (rc != 0) return rc;
    timer_begin(ws, "backward");
    k_backward_stored<<<ws->cus * kBucketWavesPerCu, 64, 0, ws->stream>>>(a);
    GS_CHECK_LAUNCH("k_backward_stored");
    timer_end(ws);
    return 0;
  }
  if (a.persistent != 0u) {
    const int rc = fill_u32(ws, a.counter, 1, 0u);
    if (rc != 0) return rc;
  }
  timer_begin(ws, "backward");
  if (ws->shape.deterministic) {
    k_backward<true><<<grid, kThreads, 0, ws->stream>>>(a);
    GS_CHECK_LAUNCH("k_backward");
    const uint32_t blocks = std::min<uint32_t>(div_up(uint64_t(ws->shape.max_capacity) * kGrad, 256), 4096);
    k_grad_from_fixed<<<dim3(std::max<uint32_t>(blocks, 1), count), 256, 0, ws->stream>>>(ws->dev_views);
    GS_CHECK_LAUNCH("k_grad_from_fixed");
  } else {
    k_backward<false><<<grid, kThreads, 0, ws->stream>>>(a);
    GS_CHECK_LAUNCH("k_backward");
  }
  timer_end(ws);
  return 0;
}

}  // namespace gsh

extern "C" {

int gs_hip_forward(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "forward: null argument");
  const int rc = gsh::upload_table(ws, f, 1);
  return rc != 0 ? rc : gsh::launch_forward(ws, 1);
}

int gs_hip_backward(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "backward: null argument");
  const int rc = gsh::upload_table(ws, f, 1);
  return rc != 0 ? rc : gsh::launch_backward(ws, 1);
}

}  // extern "C"
