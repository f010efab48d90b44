// Tile rasterizer: front-to-back compositing (forward) and the
// recompute-transmittance adjoint (backward), one 16x16 tile per workgroup of
// 256 threads.  Each wave owns an 8x8 quadrant of the tile.
//
// Both kernels stage the tile's splats through LDS in batches of 256 (one
// coalesced 48 B record load per thread), and while loading each thread tests
// its splat's footprint bbox against the four quadrants: a ballot per quadrant
// gives each wave a bitmask of the batch's splats that can reach its pixels,
// and the wave walks only those bits.  Splat parameters are then LDS broadcast
// reads, and the per-pixel work is pure VALU.
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
// per pixel is stored by the forward besides T, LAST and dL/dC.  The nine
// per-splat gradients are summed over each wave with DPP/permlane, over the
// four waves in LDS (ds_add_f32), and added to the slot's planar GRAD2D with
// one global float atomic each.
#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;

constexpr int kThreads = 256;
constexpr int kBatch = 256;

struct RasterArgs {
  const uint2* __restrict__ ranges;
  const uint32_t* __restrict__ tile_order;  // launch order: longest tile lists first
  const uint32_t* __restrict__ pair_gaussians;
  const float4* __restrict__ records;
  float4* __restrict__ image;
  uint32_t* __restrict__ last;
  const float4* __restrict__ dldc;
  float* __restrict__ grad2d;
  uint32_t capacity;
};

struct PixelCoord {
  int px, py;      // pixel
  int rx0, ry0;    // the wave's quadrant origin
  bool inside;
  uint32_t tile;
  int wave, lane;
};

__device__ __forceinline__ PixelCoord pixel_coord(const gs_hip_frame& f, uint32_t tile) {
  PixelCoord c;
  c.tile = tile;
  const int tx = int(c.tile % f.tiles_x), ty = int(c.tile / f.tiles_x);
  c.wave = threadIdx.x >> 6;
  c.lane = threadIdx.x & 63;
  c.rx0 = tx * kTile + (c.wave & 1) * 8;
  c.ry0 = ty * kTile + (c.wave >> 1) * 8;
  c.px = c.rx0 + (c.lane & 7);
  c.py = c.ry0 + (c.lane >> 3);
  c.inside = c.px < int(f.cam.width) && c.py < int(f.cam.height);
  return c;
}

// Power and alpha of a splat at a pixel centre, shared by both kernels so
// they make the same skip decisions on the same bits.
struct SplatEval {
  float dx, dy, power, G, raw, alpha;
};

__device__ __forceinline__ SplatEval eval_splat(float4 sa, float4 sb, float cx, float cy) {
  SplatEval e;
  e.dx = sa.x - cx;
  e.dy = sa.y - cy;
  e.power = -0.5f * (sa.z * e.dx * e.dx + sb.x * e.dy * e.dy) - sa.w * e.dx * e.dy;
  e.G = __expf(e.power);
  e.raw = sb.y * e.G;
  e.alpha = fminf(kMaxAlpha, e.raw);
  return e;
}

__global__ void __launch_bounds__(kThreads) k_forward(gs_hip_frame f, RasterArgs a) {
  __shared__ float4 s_a[kBatch], s_b[kBatch], s_c[kBatch];
  __shared__ uint64_t s_mask[4][kBatch / 64];
  const PixelCoord pc = pixel_coord(f, a.tile_order[blockIdx.x]);
  const uint2 range = a.ranges[pc.tile];
  const float cx = float(pc.px) + 0.5f, cy = float(pc.py) + 0.5f;
  bool done = !pc.inside;
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  uint32_t last = 0u;
  const int tx0 = (pc.tile % f.tiles_x) * kTile, ty0 = (pc.tile / f.tiles_x) * kTile;

  for (uint32_t start = range.x; start < range.y; start += kBatch) {
    if (__syncthreads_and(done)) break;
    const uint32_t n = min(uint32_t(kBatch), range.y - start);
    bool hit[4] = {false, false, false, false};
    if (threadIdx.x < n) {
      const uint32_t g = a.pair_gaussians[start + threadIdx.x];
      const float4* r = a.records + size_t(g) * 3u;
      const float4 ra = r[0], rb = r[1], rc = r[2];
      s_a[threadIdx.x] = ra;
      s_b[threadIdx.x] = rb;
      s_c[threadIdx.x] = rc;
      const uint32_t lo = __float_as_uint(rc.y), hi = __float_as_uint(rc.z);
      for (int q = 0; q < 4; ++q) {
        const int qx = tx0 + (q & 1) * 8, qy = ty0 + (q >> 1) * 8;
        hit[q] = bbox_hits(lo, hi, qx, qy, qx + 7, qy + 7);
      }
    }
    for (int q = 0; q < 4; ++q) {
      const uint64_t m = __ballot(hit[q]);
      if (pc.lane == 0) s_mask[q][pc.wave] = m;
    }
    __syncthreads();
    bool wave_live = __ballot(!done) != 0ull;
    for (int chunk = 0; chunk < kBatch / 64 && wave_live; ++chunk) {
      uint64_t m = s_mask[pc.wave][chunk];
      while (m != 0ull) {
        const int bit = __builtin_ctzll(m);
        m &= m - 1ull;
        const int j = chunk * 64 + bit;
        if (!done) {
          const SplatEval e = eval_splat(s_a[j], s_b[j], cx, cy);
          if (e.power <= 0.0f) {
            const float alpha = e.alpha;
            if (alpha >= kMinAlpha) {
              const float next_T = T * (1.0f - alpha);
              if (next_T < kMinTransmittance) {
                done = true;
              } else {
                const float4 sc = s_c[j];
                const float w = alpha * T;
                const float4 sb = s_b[j];
                cr += sb.z * w;
                cg += sb.w * w;
                cb += sc.x * w;
                T = next_T;
                last = start + uint32_t(j) - range.x + 1u;
              }
            }
          }
        }
        if (__ballot(!done) == 0ull) {
          wave_live = false;
          break;
        }
      }
    }
    __syncthreads();
  }
  if (pc.inside) {
    const uint32_t pix = uint32_t(pc.py) * f.cam.width + uint32_t(pc.px);
    a.image[pix] = make_float4(cr, cg, cb, T);
    a.last[pix] = last;
  }
}

__global__ void __launch_bounds__(kThreads) k_backward(gs_hip_frame f, RasterArgs a) {
  __shared__ float4 s_a[kBatch], s_b[kBatch], s_c[kBatch];
  __shared__ uint32_t s_g[kBatch];
  __shared__ float s_grad[GS_HIP_GRAD2D_FLOATS][kBatch];
  __shared__ uint64_t s_mask[4][kBatch / 64];
  __shared__ uint32_t s_wave_last[4];
  const PixelCoord pc = pixel_coord(f, a.tile_order[blockIdx.x]);
  const uint2 range = a.ranges[pc.tile];
  const float cx = float(pc.px) + 0.5f, cy = float(pc.py) + 0.5f;
  const int tx0 = (pc.tile % f.tiles_x) * kTile, ty0 = (pc.tile / f.tiles_x) * kTile;

  float T_final = 0.0f, gr = 0.0f, gg = 0.0f, gb = 0.0f;
  uint32_t last = 0u;
  if (pc.inside) {
    const uint32_t pix = uint32_t(pc.py) * f.cam.width + uint32_t(pc.px);
    T_final = a.image[pix].w;
    last = a.last[pix];
    const float4 d = a.dldc[pix];
    gr = d.x, gg = d.y, gb = d.z;
  }
  const float bg_dot = f.background[0] * gr + f.background[1] * gg + f.background[2] * gb;
  // Back-to-front state: T in front of the current splat, colour behind it.
  float T = T_final, acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;
  float last_alpha = 0.0f, last_r = 0.0f, last_g = 0.0f, last_b = 0.0f;

  const uint32_t wl = wave_max_u32(last);
  if (pc.lane == 0) s_wave_last[pc.wave] = wl;
  __syncthreads();
  const uint32_t tile_last = max(max(s_wave_last[0], s_wave_last[1]), max(s_wave_last[2], s_wave_last[3]));
  const uint32_t end = range.x + tile_last;

  for (uint32_t bend = end; bend > range.x;) {
    const uint32_t bstart = bend > range.x + kBatch ? bend - kBatch : range.x;
    const uint32_t n = bend - bstart;
    bool hit[4] = {false, false, false, false};
    if (threadIdx.x < n) {
      const uint32_t k = bstart + threadIdx.x;
      const uint32_t g = a.pair_gaussians[k];
      const float4* r = a.records + size_t(g) * 3u;
      const float4 ra = r[0], rb = r[1], rc = r[2];
      s_a[threadIdx.x] = ra;
      s_b[threadIdx.x] = rb;
      s_c[threadIdx.x] = rc;
      s_g[threadIdx.x] = g;
      for (int c = 0; c < GS_HIP_GRAD2D_FLOATS; ++c) s_grad[c][threadIdx.x] = 0.0f;
      const uint32_t lo = __float_as_uint(rc.y), hi = __float_as_uint(rc.z);
      const uint32_t pos = k - range.x;
      for (int q = 0; q < 4; ++q) {
        const int qx = tx0 + (q & 1) * 8, qy = ty0 + (q >> 1) * 8;
        hit[q] = pos < s_wave_last[q] && bbox_hits(lo, hi, qx, qy, qx + 7, qy + 7);
      }
    }
    for (int q = 0; q < 4; ++q) {
      const uint64_t m = __ballot(hit[q]);
      if (pc.lane == 0) s_mask[q][pc.wave] = m;
    }
    __syncthreads();
    for (int chunk = kBatch / 64 - 1; chunk >= 0; --chunk) {
      uint64_t m = s_mask[pc.wave][chunk];
      while (m != 0ull) {
        const int bit = 63 - __builtin_clzll(m);
        m &= ~(1ull << bit);
        const int j = chunk * 64 + bit;
        const uint32_t pos = bstart + uint32_t(j) - range.x;
        float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f, g4 = 0.0f, g5 = 0.0f, g6 = 0.0f, g7 = 0.0f, g8 = 0.0f;
        bool contributes = false;
        if (pos < last) {
          const float4 sa = s_a[j];
          const float4 sb = s_b[j];
          const SplatEval e = eval_splat(sa, sb, cx, cy);
          const float dx = e.dx, dy = e.dy, G = e.G, raw = e.raw, alpha = e.alpha;
          if (e.power <= 0.0f && alpha >= kMinAlpha) {
            contributes = true;
            const float4 sc = s_c[j];
            const float one_minus = 1.0f - alpha;
            T = T / one_minus;
            acc_r = last_alpha * last_r + (1.0f - last_alpha) * acc_r;
            acc_g = last_alpha * last_g + (1.0f - last_alpha) * acc_g;
            acc_b = last_alpha * last_b + (1.0f - last_alpha) * acc_b;
            last_r = sb.z, last_g = sb.w, last_b = sc.x;
            float dL_dalpha = ((sb.z - acc_r) * gr + (sb.w - acc_g) * gg + (sc.x - acc_b) * gb) * T;
            last_alpha = alpha;
            dL_dalpha += (-T_final / one_minus) * bg_dot;
            const float wT = alpha * T;
            g6 = wT * gr, g7 = wT * gg, g8 = wT * gb;
            if (raw < kMaxAlpha) {
              g5 = G * dL_dalpha;
              const float dpow = raw * dL_dalpha;
              g0 = dpow * -(sa.z * dx + sa.w * dy);
              g1 = dpow * -(sb.x * dy + sa.w * dx);
              g2 = dpow * (-0.5f * dx * dx);
              g3 = dpow * (-dx * dy);
              g4 = dpow * (-0.5f * dy * dy);
            }
          }
        }
        if (__ballot(contributes) != 0ull) {
          g0 = wave_sum(g0);
          g1 = wave_sum(g1);
          g2 = wave_sum(g2);
          g3 = wave_sum(g3);
          g4 = wave_sum(g4);
          g5 = wave_sum(g5);
          g6 = wave_sum(g6);
          g7 = wave_sum(g7);
          g8 = wave_sum(g8);
          if (pc.lane == 0) {
            atomicAdd(&s_grad[0][j], g0);
            atomicAdd(&s_grad[1][j], g1);
            atomicAdd(&s_grad[2][j], g2);
            atomicAdd(&s_grad[3][j], g3);
            atomicAdd(&s_grad[4][j], g4);
            atomicAdd(&s_grad[5][j], g5);
            atomicAdd(&s_grad[6][j], g6);
            atomicAdd(&s_grad[7][j], g7);
            atomicAdd(&s_grad[8][j], g8);
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < n) {
      const uint32_t g = s_g[threadIdx.x];
      for (int c = 0; c < GS_HIP_GRAD2D_FLOATS; ++c) {
        const float v = s_grad[c][threadIdx.x];
        if (v != 0.0f) unsafeAtomicAdd(&a.grad2d[size_t(c) * a.capacity + g], v);
      }
    }
    __syncthreads();
    bend = bstart;
  }
}

RasterArgs make_args(gs_hip_workspace* ws, gsh::View& v) {
  RasterArgs a;
  a.ranges = reinterpret_cast<const uint2*>(v.ranges);
  a.tile_order = v.tile_order;
  a.pair_gaussians = v.pair_gaussians;
  a.records = reinterpret_cast<const float4*>(v.records);
  a.image = reinterpret_cast<float4*>(v.image);
  a.last = v.last;
  a.dldc = reinterpret_cast<const float4*>(v.dldc);
  a.grad2d = v.grad2d;
  a.capacity = ws->capacity;
  return a;
}

}  // namespace

namespace gsh {

int launch_forward(gs_hip_workspace* ws, View& v, const gs_hip_frame* f) {
  const uint32_t tiles = f->tiles_x * f->tiles_y;
  timer_begin(ws, v, "forward");
  k_forward<<<tiles, kThreads, 0, v.stream>>>(*f, make_args(ws, v));
  GS_CHECK_LAUNCH("k_forward");
  timer_end(ws, v);
  return 0;
}

int launch_backward(gs_hip_workspace* ws, View& v, const gs_hip_frame* f) {
  const uint32_t tiles = f->tiles_x * f->tiles_y;
  timer_begin(ws, v, "backward");
  k_backward<<<tiles, kThreads, 0, v.stream>>>(*f, make_args(ws, v));
  GS_CHECK_LAUNCH("k_backward");
  timer_end(ws, v);
  return 0;
}

}  // namespace gsh

extern "C" {

int gs_hip_forward(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "forward: null argument");
  return gsh::launch_forward(ws, gsh::view0(ws), f);
}

int gs_hip_backward(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "backward: null argument");
  return gsh::launch_backward(ws, gsh::view0(ws), f);
}

}  // extern "C"
