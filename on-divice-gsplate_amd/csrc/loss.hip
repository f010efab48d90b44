// Loss: (1 - lambda) L1 + lambda (1 - SSIM) between the render
// x = C + T bg and the target y, per channel, 11x11 Gaussian window
// (sigma 1.5), zero padding outside the image (engine/shaders/train_ssim.comp,
// train_ssim_grad.comp, train_forward.comp:119-140).
//
// k_ssim: one 16x16 tile per workgroup; stages the 26x26 halo of x and y of
// one channel in LDS, blurs x, y, x^2, y^2, xy along rows then columns, and
// writes per pixel and channel the parts of the SSIM gradient
//   A = dSSIM/dmu_x (through the variances too), B = dSSIM/dE_xx, C = dSSIM/dE_xy
// plus the tile's sums of |e|, e^2 and SSIM.
// k_loss_grad: blurs A, B, C back with the same window and forms
//   dL/dx_q = (1 - lambda) sign(e_q) / (3HW)
//             - lambda / (3HW) ((G*A)(q) + 2 x_q (G*B)(q) + y_q (G*C)(q)).
#include <mutex>
#include <string>
#include <vector>

#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;

constexpr int kHalo = kTile + 2 * kSsimRadius;  // 26
// LDS row stride of the halo: 48 = 16 mod 32, so the two 16-pixel rows each
// half-wave reads in the horizontal blur fall on disjoint banks (with 26 they
// overlapped: 1.1-1.7 bank-conflict cycles per LDS instruction).
constexpr int kStride = 48;

struct LossArgs {
  const float4* __restrict__ image;
  const uint32_t* __restrict__ target;
  float* __restrict__ maps;  // [9][pixels]
  float4* __restrict__ dldc;
  float4* __restrict__ tile_loss;
};

// Render x = C + T bg and target y of a pixel from its IMAGE texel and RGBA8
// photo: the photo times exposure[] with GS_HIP_FLAG_EXPOSURE, over the
// background by its alpha with GS_HIP_FLAG_MASKED_TARGET.
__device__ __forceinline__ void xy_of(const gs_hip_frame& f, float4 im, uint32_t t, float rgb_x[3], float rgb_y[3]) {
  rgb_x[0] = im.x + im.w * f.background[0];
  rgb_x[1] = im.y + im.w * f.background[1];
  rgb_x[2] = im.z + im.w * f.background[2];
  float tr = float(t & 0xffu) * (1.0f / 255.0f), tg = float((t >> 8) & 0xffu) * (1.0f / 255.0f);
  float tb = float((t >> 16) & 0xffu) * (1.0f / 255.0f);
  const float ta = float(t >> 24) * (1.0f / 255.0f);
  if ((f.flags & GS_HIP_FLAG_EXPOSURE) != 0u) {  // the photo at the model's brightness
    tr *= f.exposure[0];
    tg *= f.exposure[1];
    tb *= f.exposure[2];
  }
  if ((f.flags & GS_HIP_FLAG_MASKED_TARGET) != 0u) {
    rgb_y[0] = f.background[0] * (1.0f - ta) + tr * ta;
    rgb_y[1] = f.background[1] * (1.0f - ta) + tg * ta;
    rgb_y[2] = f.background[2] * (1.0f - ta) + tb * ta;
  } else {
    rgb_y[0] = tr, rgb_y[1] = tg, rgb_y[2] = tb;
  }
}

// Halo elements per thread: the 26 x 26 halo over 256 threads.
constexpr int kHaloIters = (kHalo * kHalo + 255) / 256;

// Halo element e of the tile whose halo starts at (ox, oy): its pixel index,
// or -1 outside the image (and past the halo).
__device__ __forceinline__ int halo_pixel(const gs_hip_frame& f, int e, int ox, int oy) {
  const int gx = ox + e % kHalo, gy = oy + e / kHalo;
  const bool in = e < kHalo * kHalo && gx >= 0 && gy >= 0 && gx < int(f.cam.width) && gy < int(f.cam.height);
  return in ? gy * int(f.cam.width) + gx : -1;
}

// Tile of block b among n (a bijection on [0, n)): consecutive blocks go to
// the 8 XCDs in turn, so XCD x gets the contiguous run of tiles
// [x q + min(x, r), ...) (q = n / 8, r = n % 8) and neighbouring tiles' halos
// meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t x = b & 7u, idx = b >> 3, q = n >> 3, r = n & 7u;
  return x * q + min(x, r) + idx;
}

__device__ __forceinline__ LossArgs loss_args(const gsh::ViewDev& v) {
  LossArgs a;
  a.image = v.image;
  a.target = v.target;
  a.maps = v.maps;
  a.dldc = v.dldc;
  a.tile_loss = v.tile_loss;
  return a;
}

// The exposure gain's least-squares sums over one tile (train_forward.comp:142-151):
// where the photo is opaque (alpha > 0.95) and the render nearly so (T < 0.05),
// C^2 and C times the photo as stored (C the model's colour before the
// background), those pixels, and the pixels with any alpha.  pix < 0: outside
// the image.  Every thread of the block calls it.
__device__ void exposure_fit(const gsh::ViewDev& view, uint32_t tile, int pix) {
  __shared__ float s_fit[8][4];
  float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (pix >= 0) {
    const float4 im = gload(view.image + pix);
    const uint32_t t = gmem(view.target)[pix];
    const float ta = float(t >> 24) * (1.0f / 255.0f);
    if (ta > 0.95f && im.w < 0.05f) {
      const float photo[3] = {float(t & 0xffu) * (1.0f / 255.0f), float((t >> 8) & 0xffu) * (1.0f / 255.0f),
                              float((t >> 16) & 0xffu) * (1.0f / 255.0f)};
      const float C[3] = {im.x, im.y, im.z};
      for (int c = 0; c < 3; ++c) {
        v[c] = C[c] * C[c];
        v[4 + c] = C[c] * photo[c];
      }
      v[3] = 1.0f;
    }
    v[7] = ta > 0.0f ? 1.0f : 0.0f;
  }
  const int wave = threadIdx.x >> 6;
  for (int k = 0; k < 8; ++k) {
    const float w = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) s_fit[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    float r[4];
    for (int k = 0; k < 4; ++k) {
      const int q = int(threadIdx.x) * 4 + k;
      r[k] = s_fit[q][0] + s_fit[q][1] + s_fit[q][2] + s_fit[q][3];
    }
    gstore(view.tile_fit + size_t(tile) * 2u + threadIdx.x, make_float4(r[0], r[1], r[2], r[3]));
  }
}

// One tile (blockIdx.x) of one view (blockIdx.y).
__global__ void __launch_bounds__(256) k_ssim(const gsh::ViewDev* __restrict__ views) {
  const gsh::ViewDev& view = views[blockIdx.y];
  // A smaller view of the launch leaves its extra blocks idle; its own tiles
  // are dealt over all 8 XCDs (with the launch's width they would all fall
  // on the first XCDs' runs).
  if (blockIdx.x >= view.tile_count) return;  // the whole block leaves
  const uint32_t tile = xcd_tile(blockIdx.x, view.tile_count);
  const gs_hip_frame& f = view.f;
  const LossArgs a = loss_args(view);
  __shared__ float s_x[3][kHalo * kStride];
  __shared__ float s_y[3][kHalo * kStride];
  __shared__ float s_h[5][kHalo * kTile];
  __shared__ float s_sum[3][4];
  const int tx = int(tile % f.tiles_x), ty = int(tile / f.tiles_x);
  const int ox = tx * kTile - kSsimRadius, oy = ty * kTile - kSsimRadius;
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int px = tx * kTile + lx, py = ty * kTile + ly;
  const bool inside = px < int(f.cam.width) && py < int(f.cam.height);
  const uint32_t pixels = f.cam.width * f.cam.height;
  {
    // Every halo texel this thread stages is requested before any is used.
    float4 im[kHaloIters];
    uint32_t tg[kHaloIters];
    int id[kHaloIters];
#pragma unroll
    for (int it = 0; it < kHaloIters; ++it) {
      id[it] = halo_pixel(f, int(threadIdx.x) + it * 256, ox, oy);
      im[it] = id[it] >= 0 ? gload(a.image + id[it]) : make_float4(0, 0, 0, 0);
      tg[it] = id[it] >= 0 ? gmem(a.target)[id[it]] : 0u;
    }
#pragma unroll
    for (int it = 0; it < kHaloIters; ++it) {
      const int e = int(threadIdx.x) + it * 256;
      if (e >= kHalo * kHalo) break;
      float vx[3] = {0, 0, 0}, vy[3] = {0, 0, 0};
      if (id[it] >= 0) xy_of(f, im[it], tg[it], vx, vy);
      const int at = (e / kHalo) * kStride + e % kHalo;
      for (int c = 0; c < 3; ++c) {
        s_x[c][at] = vx[c];
        s_y[c][at] = vy[c];
      }
    }
  }
  __syncthreads();
  const float c1 = f.ssim_c1, c2 = f.ssim_c2;
  float abs_e = 0.0f, sq_e = 0.0f, ssim_sum = 0.0f;
  const uint32_t pix = uint32_t(py) * f.cam.width + uint32_t(px);
  for (int c = 0; c < 3; ++c) {
    for (int e = threadIdx.x; e < kHalo * kTile; e += 256) {
      const int row = e / kTile, col = e % kTile;
      // Horizontal pass of the five moments x, y, xx, yy, xy.
      float sum_x = 0.0f, sum_y = 0.0f, sum_xx = 0.0f, sum_yy = 0.0f, sum_xy = 0.0f;
      const float* row_x = &s_x[c][row * kStride + col];
      const float* row_y = &s_y[c][row * kStride + col];
      for (int k = 0; k <= 2 * kSsimRadius; ++k) {
        const float w = kSsimWindow[k], x = row_x[k], y = row_y[k], wx = w * x, wy = w * y;
        sum_x += w * x, sum_y += w * y;
        sum_xx += wx * x, sum_yy += wy * y, sum_xy += wx * y;
      }
      const float sums[5] = {sum_x, sum_y, sum_xx, sum_yy, sum_xy};
      for (int m = 0; m < 5; ++m) s_h[m][e] = sums[m];
    }
    __syncthreads();
    float mx = 0.0f, my = 0.0f, exx = 0.0f, eyy = 0.0f, exy = 0.0f;
    for (int k = 0; k <= 2 * kSsimRadius; ++k) {
      const float w = kSsimWindow[k];
      const int e = (ly + k) * kTile + lx;
      mx += w * s_h[0][e];
      my += w * s_h[1][e];
      exx += w * s_h[2][e];
      eyy += w * s_h[3][e];
      exy += w * s_h[4][e];
    }
    const float vx = exx - mx * mx, vy = eyy - my * my, cxy = exy - mx * my;
    const float l = 2.0f * mx * my + c1, cs = 2.0f * cxy + c2;
    const float m = mx * mx + my * my + c1, v = vx + vy + c2;
    const float inv = 1.0f / (m * v);
    const float ssim = l * cs * inv;
    const float d_cxy = 2.0f * l * inv;  // dSSIM/dE_xy
    const float d_vx = -ssim / v;        // dSSIM/dE_xx
    const float A = 2.0f * my * cs * inv - ssim * 2.0f * mx / m - 2.0f * mx * d_vx - my * d_cxy;
    if (inside) {
      gmem(a.maps)[size_t(c) * pixels + pix] = A;
      gmem(a.maps)[size_t(3 + c) * pixels + pix] = d_vx;
      gmem(a.maps)[size_t(6 + c) * pixels + pix] = d_cxy;
      const int centre = (ly + kSsimRadius) * kStride + lx + kSsimRadius;
      const float err = s_x[c][centre] - s_y[c][centre];
      abs_e += fabsf(err);
      sq_e += err * err;
      ssim_sum += ssim;
    }
    __syncthreads();
  }
  abs_e = wave_sum(abs_e);
  sq_e = wave_sum(sq_e);
  ssim_sum = wave_sum(ssim_sum);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_sum[0][wave] = abs_e;
    s_sum[1][wave] = sq_e;
    s_sum[2][wave] = ssim_sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float r[3];
    for (int q = 0; q < 3; ++q) r[q] = s_sum[q][0] + s_sum[q][1] + s_sum[q][2] + s_sum[q][3];
    gstore(a.tile_loss + tile, make_float4(r[0], r[1], r[2], 0.0f));
  }
  if ((f.flags & GS_HIP_FLAG_EXPOSURE) != 0u) exposure_fit(view, tile, inside ? int(pix) : -1);
}

__global__ void __launch_bounds__(256) k_loss_grad(const gsh::ViewDev* __restrict__ views) {
  const gsh::ViewDev& view = views[blockIdx.y];
  // A smaller view of the launch leaves its extra blocks idle; its own tiles
  // are dealt over all 8 XCDs (with the launch's width they would all fall
  // on the first XCDs' runs).
  if (blockIdx.x >= view.tile_count) return;  // the whole block leaves
  const uint32_t tile = xcd_tile(blockIdx.x, view.tile_count);
  const gs_hip_frame& f = view.f;
  const LossArgs a = loss_args(view);
  __shared__ float s_m[3][kHalo * kStride];
  __shared__ float s_h[3][kHalo * kTile];
  const int tx = int(tile % f.tiles_x), ty = int(tile / f.tiles_x);
  const int ox = tx * kTile - kSsimRadius, oy = ty * kTile - kSsimRadius;
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int px = tx * kTile + lx, py = ty * kTile + ly;
  const bool inside = px < int(f.cam.width) && py < int(f.cam.height);
  const uint32_t pixels = f.cam.width * f.cam.height;
  const uint32_t pix = uint32_t(py) * f.cam.width + uint32_t(px);
  float x[3] = {0, 0, 0}, y[3] = {0, 0, 0};
  if (inside) xy_of(f, gload(a.image + pix), gmem(a.target)[pix], x, y);
  const float scale = 1.0f / (3.0f * float(pixels));
  const float lambda = f.lambda_dssim;
  float grad[3];
  for (int c = 0; c < 3; ++c) {
    const float e = x[c] - y[c];
    const float sgn = e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f);
    grad[c] = sgn * ((1.0f - lambda) * scale);
  }
  if (lambda != 0.0f) {
    int id[kHaloIters];
#pragma unroll
    for (int it = 0; it < kHaloIters; ++it) id[it] = halo_pixel(f, int(threadIdx.x) + it * 256, ox, oy);
    // A channel's nine halo values per thread, all requested at once; the
    // next channel's are requested before this channel's blurs, so their
    // latency passes under that work.
    float mv[kHaloIters][3];
    auto request = [&](int c) {
#pragma unroll
      for (int it = 0; it < kHaloIters; ++it) {
        for (int q = 0; q < 3; ++q) {
          mv[it][q] = id[it] >= 0 ? gmem(a.maps)[size_t(3 * q + c) * pixels + id[it]] : 0.0f;
        }
      }
    };
    request(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int it = 0; it < kHaloIters; ++it) {
        const int e = int(threadIdx.x) + it * 256;
        if (e >= kHalo * kHalo) break;
        const int at = (e / kHalo) * kStride + e % kHalo;
        for (int q = 0; q < 3; ++q) s_m[q][at] = mv[it][q];
      }
      __syncthreads();
      if (c + 1 < 3) request(c + 1);
      for (int e = threadIdx.x; e < kHalo * kTile; e += 256) {
        const int row = e / kTile, col = e % kTile;
        for (int q = 0; q < 3; ++q) {
          float sum = 0.0f;
          for (int k = 0; k <= 2 * kSsimRadius; ++k) sum += kSsimWindow[k] * s_m[q][row * kStride + col + k];
          s_h[q][e] = sum;
        }
      }
      __syncthreads();
      float g[3] = {0.0f, 0.0f, 0.0f};
      for (int q = 0; q < 3; ++q) {
        for (int k = 0; k <= 2 * kSsimRadius; ++k) g[q] += kSsimWindow[k] * s_h[q][(ly + k) * kTile + lx];
      }
      grad[c] -= (lambda * scale) * (g[0] + 2.0f * x[c] * g[1] + y[c] * g[2]);
      __syncthreads();
    }
  }
  if (inside) gstore(a.dldc + pix, make_float4(grad[0], grad[1], grad[2], 0.0f));
}

}  // namespace

namespace gsh {

int launch_loss(gs_hip_workspace* ws, uint32_t count) {
  const dim3 grid(ws->shape.max_tiles, count);
  timer_begin(ws, "ssim");
  k_ssim<<<grid, 256, 0, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_ssim");
  timer_end(ws);
  timer_begin(ws, "loss_grad");
  k_loss_grad<<<grid, 256, 0, ws->stream>>>(ws->dev_views);
  GS_CHECK_LAUNCH("k_loss_grad");
  timer_end(ws);
  return 0;
}

int launch_readback(gs_hip_workspace* ws, uint32_t count, bool losses) {
  GS_HIP_TRY(hipMemcpyAsync(ws->host_totals, ws->totals_all, size_t(count) * 64, hipMemcpyDeviceToHost,
                            ws->stream));
  if (losses) {
    GS_HIP_TRY(hipMemcpyAsync(ws->host_tile_loss, ws->tile_loss_all, size_t(ws->shape.total_tiles) * 16,
                              hipMemcpyDeviceToHost, ws->stream));
  }
  return 0;
}

void sum_losses(gs_hip_workspace* ws, uint32_t v, double losses[3]) {
  const uint32_t tiles = v + 1 < ws->tile_base.size() ? ws->tile_base[v + 1] - ws->tile_base[v]
                                                       : ws->shape.total_tiles - ws->tile_base[v];
  const float* t = ws->host_tile_loss + size_t(ws->tile_base[v]) * 4;
  double s[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = 0; i < tiles; ++i) {
    for (int q = 0; q < 3; ++q) s[q] += t[i * 4 + q];
  }
  for (int q = 0; q < 3; ++q) losses[q] = s[q];
}

}  // namespace gsh

extern "C" {

int gs_hip_loss(gs_hip_workspace* ws, const gs_hip_frame* f, double losses[3]) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "loss: null argument");
  int rc = gsh::upload_table(ws, f, 1);
  if (rc == 0) rc = gsh::launch_loss(ws, 1);
  if (rc != 0) return rc;
  if (losses != nullptr) {
    rc = gsh::launch_readback(ws, 1, true);
    if (rc != 0) return rc;
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
    gsh::sum_losses(ws, 0, losses);
  }
  return 0;
}

int gs_hip_exposure_sums(gs_hip_workspace* ws, uint32_t view, double sums[8]) {
  if (ws == nullptr || sums == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "exposure_sums: null argument");
  if (view >= ws->views.size()) return gsh::set_error(GS_HIP_E_ARGUMENT, "exposure_sums: no such view");
  const gsh::View& v = *ws->views[view];
  std::vector<float> fit(size_t(v.tile_grid) * 8);
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  GS_HIP_TRY(hipMemcpy(fit.data(), v.tile_fit, fit.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int k = 0; k < 8; ++k) sums[k] = 0.0;
  for (uint32_t t = 0; t < v.tile_grid; ++t) {
    for (int k = 0; k < 8; ++k) sums[k] += fit[size_t(t) * 8 + k];
  }
  return 0;
}

}  // extern "C"

// ---- the loss without allocation -------------------------------------------------
// gs_hip_loss_noalloc runs k_ssim and k_loss_grad on caller-owned image,
// target and dL/dC buffers; everything else they use lives in `scratch`, a
// file-scope device array each GPU holds from the library's load to the
// process's end: [0, 1 KiB) the one-view table, then the three loss sums
// (double), the per-tile loss and exposure-fit partial sums, and the nine
// SSIM gradient maps.  Calls that overlap on one GPU share it and race.
namespace {

constexpr uint32_t kNoAllocMaxPixels = 1u << 21;  // e.g. 1920 x 1080 = 2,073,600
constexpr uint32_t kNoAllocMaxTiles = 1u << 14;   // 1920 x 1080: 120 x 68 = 8160
constexpr size_t kNoAllocSums = 1024;             // double[3] after the table
constexpr size_t kNoAllocTileLoss = 4096;
constexpr size_t kNoAllocTileFit = kNoAllocTileLoss + size_t(kNoAllocMaxTiles) * 16;
constexpr size_t kNoAllocMaps = kNoAllocTileFit + size_t(kNoAllocMaxTiles) * 32;
constexpr size_t kNoAllocBytes = kNoAllocMaps + size_t(9) * kNoAllocMaxPixels * 4;  // ~72.8 MiB
static_assert(sizeof(gsh::ViewDev) <= kNoAllocSums, "the view table must fit before the sums");

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The device address of this GPU's `scratch`, looked up on its first call
// with that device current (the host-side name is not a device address).
std::mutex g_scratch_mutex;
void* g_scratch_ptr[64] = {};

int scratch_address(void** out) {
  int dev = 0;
  GS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return gsh::set_error(GS_HIP_E_STATE, "loss_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || p == nullptr) {
      return gsh::set_error(GS_HIP_E_HIP, std::string("loss_noalloc: hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(rc));
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Writes the one-view table into scratch (the loss kernels read their view
// from device memory; a kernel writes it so the call stays capturable).
__global__ void k_put_view(gsh::ViewDev v, gsh::ViewDev* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// The three loss sums over the tiles, in tile order, in double (the order
// gs_hip_loss sums them on the host).
__global__ void k_sum_tile_losses(const float4* __restrict__ tile_loss, uint32_t tiles, double* __restrict__ sums) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t t = 0; t < tiles; ++t) {
    const float4 v = tile_loss[t];
    s0 += v.x;
    s1 += v.y;
    s2 += v.z;
  }
  sums[0] = s0, sums[1] = s1, sums[2] = s2;
}

}  // namespace

extern "C" int gs_hip_loss_noalloc(const gs_hip_frame* f, const void* image, const void* target, void* dldc,
                                   void* stream, double losses[3]) {
  gsh::clear_error();
  if (f == nullptr || image == nullptr || target == nullptr || dldc == nullptr) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "loss_noalloc: null argument");
  }
  const uint32_t w = f->cam.width, h = f->cam.height;
  const uint64_t pixels = uint64_t(w) * h, tiles = uint64_t(f->tiles_x) * f->tiles_y;
  if (w == 0 || h == 0 || f->tiles_x != (w + kTile - 1) / kTile || f->tiles_y != (h + kTile - 1) / kTile) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "loss_noalloc: image size and tile grid disagree");
  }
  if (pixels > kNoAllocMaxPixels || tiles > kNoAllocMaxTiles) {
    return gsh::set_error(GS_HIP_E_CAPACITY, "loss_noalloc: " + std::to_string(w) + "x" + std::to_string(h) +
                                                 " exceeds the scratch limit of " + std::to_string(kNoAllocMaxPixels) +
                                                 " pixels and " + std::to_string(kNoAllocMaxTiles) + " tiles");
  }
  void* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  unsigned char* s = static_cast<unsigned char*>(base);
  gsh::ViewDev v{};
  v.f = *f;
  v.image = static_cast<float4*>(const_cast<void*>(image));
  v.target = static_cast<const uint32_t*>(target);
  v.dldc = static_cast<float4*>(dldc);
  v.maps = reinterpret_cast<float*>(s + kNoAllocMaps);
  v.tile_loss = reinterpret_cast<float4*>(s + kNoAllocTileLoss);
  v.tile_fit = reinterpret_cast<float4*>(s + kNoAllocTileFit);
  v.tile_count = uint32_t(tiles);
  auto* table = reinterpret_cast<gsh::ViewDev*>(s);
  double* sums = reinterpret_cast<double*>(s + kNoAllocSums);
  hipStream_t st = static_cast<hipStream_t>(stream);
  k_put_view<<<1, 64, 0, st>>>(v, table);
  GS_CHECK_LAUNCH("k_put_view");
  const dim3 grid(uint32_t(tiles), 1);
  k_ssim<<<grid, 256, 0, st>>>(table);
  GS_CHECK_LAUNCH("k_ssim");
  k_loss_grad<<<grid, 256, 0, st>>>(table);
  GS_CHECK_LAUNCH("k_loss_grad");
  if (losses != nullptr) {
    k_sum_tile_losses<<<1, 64, 0, st>>>(v.tile_loss, uint32_t(tiles), sums);
    GS_CHECK_LAUNCH("k_sum_tile_losses");
    GS_HIP_TRY(hipMemcpyAsync(losses, sums, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    GS_HIP_TRY(hipStreamSynchronize(st));
  }
  return 0;
}
