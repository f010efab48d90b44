// Viewer renderer: the image contract of the reference's gs_render / app
// viewer (engine/shaders/splat_preprocess.comp:159-291, splat.vert,
// splat.frag, engine/src/splat_renderer.cpp:243-348), rebuilt as compute on
// the training path's binning machinery instead of a raster pipeline.
//
//  * k_view_project, one thread per slot: camera transform; opacity (with
//    the optional mip-splat compensation sqrt(det_raw / det)); the 2D
//    covariance dilated by 0.3 px^2; its eigen-axes as quad half-axes
//    v_i * min(sqrt(2 lambda_i), 2048) (lambda_2 >= 0.1); on-screen test with
//    the quad extent 2 |axis1|; SH colour from coefficients stored at fp16
//    precision, clamped at 0, colour and opacity rounded to fp16 as the
//    reference packs them; a 16-bit key 65534 - floor(65534 log(z/near) /
//    log(far/near)) so that far splats sort first; the quad's pixel bbox and
//    its tile rectangle.
//  * the depth sort (16 bits) and a stable tile sort of (tile, splat) pairs
//    over each quad's tile rectangle: every tile's list in drawing order.
//  * k_view_composite, one 16x16 tile per workgroup (a wave per 8x8
//    quadrant, culled by the quad bbox): back to front from the clear
//    colour; a pixel centre inside the quad (|local| <= 2 on both axes) gets
//    alpha = min(o exp(-|local|^2), 0.99), dropped below 1/255; blending is
//    premultiplied "over" into an RGBA8 target: the fragment colour clamps to
//    [0, 1] and every blend result is rounded to 8 bits, as a UNORM
//    attachment does.  Contraction off and the library's own exp: the image
//    is bit-exact with the oracle's per-pixel restatement.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;

constexpr int kBatch = 256;

struct ViewArgs {
  const float4* __restrict__ params;
  const float* __restrict__ sh;
  const gsh::ViewDev* __restrict__ views;
  uint32_t capacity;
  uint32_t sh_coeffs;
  uint32_t antialiased;
  float clear[4];
  uint8_t* __restrict__ out;  // RGBA8 [H][W]
  uint32_t* __restrict__ visible;
};

__device__ __forceinline__ float to_half(float x) { return __half2float(__float2half_rn(x)); }

// The reference's eval_sh (splat_preprocess.comp): DC + rest over the basis,
// + 0.5, clamped at 0.
__device__ __forceinline__ void view_colour(float4 p3, const float* __restrict__ sh, uint32_t capacity, uint32_t i,
                                            int coeffs, uint32_t degree, float x, float y, float z, float c[3]) {
#pragma clang fp contract(off)
  float basis[16];
  sh_basis(x, y, z, degree, basis);
  c[0] = basis[0] * p3.x, c[1] = basis[0] * p3.y, c[2] = basis[0] * p3.z;
  for (int k = 0; k < GS_HIP_MAX_SH_REST; ++k) {
    if (k < coeffs) {
      const float* s = sh + (size_t(k) * capacity + i) * 3u;
      c[0] = c[0] + basis[k + 1] * s[0];
      c[1] = c[1] + basis[k + 1] * s[1];
      c[2] = c[2] + basis[k + 1] * s[2];
    }
  }
  for (int q = 0; q < 3; ++q) c[q] = fmaxf(c[q] + 0.5f, 0.0f);
}

__global__ void __launch_bounds__(256) k_view_project(ViewArgs a) {
#pragma clang fp contract(off)
  const gsh::ViewDev& v = a.views[0];
  const gs_hip_frame& f = v.f;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.slots) return;
  const float4 p0 = a.params[i * 4u], p1 = a.params[i * 4u + 1u];
  const float* R = f.cam.rotation;
  const float tx = R[0] * p0.x + R[1] * p0.y + R[2] * p0.z + f.cam.translation[0];
  const float ty = R[3] * p0.x + R[4] * p0.y + R[5] * p0.z + f.cam.translation[1];
  const float tz = R[6] * p0.x + R[7] * p0.y + R[8] * p0.z + f.cam.translation[2];
  float opacity = gs_sigmoid(p0.w);
  uint32_t key = 0xffffu, tiles = 0u;
  Record rec;
  rec.a = rec.b = rec.c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t rect = 0u, rect_hi = 0u;  // tile rectangle corners, x | y << 16
  if (p1.w > 0.5f && tz > kNear && opacity >= kMinAlpha) {
    const float fx = f.cam.fx, fy = f.cam.fy, cx = f.cam.cx, cy = f.cam.cy;
    const float W = float(f.cam.width), H = float(f.cam.height);
    const float inv_z = 1.0f / tz;
    const float mx = fx * tx * inv_z + cx, my = fy * ty * inv_z + cy;
    const float lim_x = 1.3f * fmaxf(cx, W - cx) / fx, lim_y = 1.3f * fmaxf(cy, H - cy) / fy;
    const float rx = fminf(fmaxf(tx * inv_z, -lim_x), lim_x), ry = fminf(fmaxf(ty * inv_z, -lim_y), lim_y);
    const float j00 = fx * inv_z, j02 = -fx * rx * inv_z, j11 = fy * inv_z, j12 = -fy * ry * inv_z;
    const float t00 = j00 * R[0] + j02 * R[6], t01 = j00 * R[1] + j02 * R[7], t02 = j00 * R[2] + j02 * R[8];
    const float t10 = j11 * R[3] + j12 * R[6], t11 = j11 * R[4] + j12 * R[7], t12 = j11 * R[5] + j12 * R[8];
    const float4 q = a.params[i * 4u + 2u];
    const float qlen = fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
    const float w = q.x / qlen, x = q.y / qlen, y = q.z / qlen, z = q.w / qlen;
    const float sx = gs_expf(p1.x), sy = gs_expf(p1.y), sz = gs_expf(p1.z);
    const float m00 = (1.0f - 2.0f * (y * y + z * z)) * sx, m01 = (2.0f * (x * y - w * z)) * sy,
                m02 = (2.0f * (x * z + w * y)) * sz;
    const float m10 = (2.0f * (x * y + w * z)) * sx, m11 = (1.0f - 2.0f * (x * x + z * z)) * sy,
                m12 = (2.0f * (y * z - w * x)) * sz;
    const float m20 = (2.0f * (x * z - w * y)) * sx, m21 = (2.0f * (y * z + w * x)) * sy,
                m22 = (1.0f - 2.0f * (x * x + y * y)) * sz;
    const float v00 = t00 * m00 + t01 * m10 + t02 * m20, v01 = t00 * m01 + t01 * m11 + t02 * m21,
                v02 = t00 * m02 + t01 * m12 + t02 * m22;
    const float v10 = t10 * m00 + t11 * m10 + t12 * m20, v11 = t10 * m01 + t11 * m11 + t12 * m21,
                v12 = t10 * m02 + t11 * m12 + t12 * m22;
    const float ra = v00 * v00 + v01 * v01 + v02 * v02;
    const float rb = v00 * v10 + v01 * v11 + v02 * v12;
    const float rc = v10 * v10 + v11 * v11 + v12 * v12;
    const float det_raw = ra * rc - rb * rb;
    const float ca = ra + kDilation, cc = rc + kDilation;
    const float det = ca * cc - rb * rb;
    if (a.antialiased != 0u) opacity = opacity * sqrtf(fmaxf(det_raw, 0.0f) / fmaxf(det, 1e-12f));
    if (det > 0.0f && opacity >= kMinAlpha) {
      const float mid = 0.5f * (ca + cc);
      const float hx = 0.5f * (ca - cc);
      const float spread = sqrtf(hx * hx + rb * rb);
      const float l1 = mid + spread, l2 = fmaxf(mid - spread, 0.1f);
      float e1x = rb, e1y = l1 - ca;
      const float elen = sqrtf(e1x * e1x + e1y * e1y);
      if (elen > 1e-12f) {
        e1x = e1x / elen, e1y = e1y / elen;
      } else {
        e1x = 1.0f, e1y = 0.0f;
      }
      const float s1 = fminf(sqrtf(2.0f * l1), 2048.0f), s2 = fminf(sqrtf(2.0f * l2), 2048.0f);
      const float a1x = e1x * s1, a1y = e1y * s1, a2x = -e1y * s2, a2y = e1x * s2;
      const float extent = 2.0f * sqrtf(a1x * a1x + a1y * a1y);
      const bool on_screen = mx + extent > 0.0f && mx - extent < W && my + extent > 0.0f && my - extent < H;
      if (on_screen) {
        const float dx = p0.x - f.cam.position[0], dy = p0.y - f.cam.position[1], dz = p0.z - f.cam.position[2];
        const float dlen = sqrtf(dx * dx + dy * dy + dz * dz);
        const int coeffs = min(sh_rest_coeffs(f.sh_degree), int(a.sh_coeffs));
        float col[3];
        view_colour(a.params[i * 4u + 3u], a.sh, a.capacity, i, coeffs, f.sh_degree, dx / dlen, dy / dlen,
                    dz / dlen, col);
        // Local quad coordinates at a pixel: d . axis / |axis|^2 (the axes are orthogonal).
        const float n1 = a1x * a1x + a1y * a1y, n2 = a2x * a2x + a2y * a2y;
        rec.a = make_float4(mx, my, a1x / n1, a1y / n1);
        rec.b = make_float4(a2x / n2, a2y / n2, to_half(opacity), to_half(col[0]));
        // Pixel bbox of the quad (corners at +-2 axis1 +-2 axis2): pixels whose centre can be inside.
        const float ex = 2.0f * (fabsf(a1x) + fabsf(a2x)), ey = 2.0f * (fabsf(a1y) + fabsf(a2y));
        const int x0 = int(fmaxf(ceilf(mx - ex - 0.5f), 0.0f)), x1 = int(fminf(floorf(mx + ex - 0.5f), W - 1.0f));
        const int y0 = int(fmaxf(ceilf(my - ey - 0.5f), 0.0f)), y1 = int(fminf(floorf(my + ey - 0.5f), H - 1.0f));
        uint32_t lo = 0u, hi = 0u;
        if (x0 <= x1 && y0 <= y1) {
          lo = uint32_t(x0) | (uint32_t(y0) << 16);
          hi = uint32_t(x1) | (uint32_t(y1) << 16);
          const uint32_t tx0 = uint32_t(x0) / kTile, tx1 = uint32_t(x1) / kTile;
          const uint32_t ty0 = uint32_t(y0) / kTile, ty1 = uint32_t(y1) / kTile;
          tiles = (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u);
          rect = tx0 | (ty0 << 16);
          rect_hi = tx1 | (ty1 << 16);
        } else {
          lo = 1u, hi = 0u;  // empty bbox: bbox_hits never true
        }
        rec.c = make_float4(to_half(col[1]), __uint_as_float(lo), __uint_as_float(hi), to_half(col[2]));
        const float u = gs_logf(tz / kNear) / gs_logf(kFar / kNear);
        key = (65534u - uint32_t(fminf(fmaxf(u, 0.0f), 1.0f) * 65534.0f)) | (tiles << 16);  // tile count rides above
        atomicAdd(a.visible, 1u);
      }
    }
  }
  float4* r = v.records + size_t(i) * 3u;
  r[0] = rec.a;
  r[1] = rec.b;
  r[2] = rec.c;
  v.footprint[size_t(i) * 2u] = make_float4(__uint_as_float(rect), __uint_as_float(rect_hi), 0.0f, 0.0f);
  v.footprint[size_t(i) * 2u + 1u] = make_float4(0.0f, 0.0f, __uint_as_float(tiles), 0.0f);
  v.tiles[i] = tiles;
  v.depth_keys[i] = key;
  v.depth_order[i] = i;
}

// Pairs (tile, splat) over each quad's tile rectangle, splats in key order.
__global__ void __launch_bounds__(256) k_view_emit(const gsh::ViewDev* __restrict__ views) {
  const gsh::ViewDev& v = views[0];
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= v.f.slots) return;
  const uint32_t i = v.depth_order[k];
  const float4 r0 = v.footprint[size_t(i) * 2u], r1 = v.footprint[size_t(i) * 2u + 1u];
  const uint32_t count = __float_as_uint(r1.z);
  if (count == 0u) return;
  const uint32_t lo = __float_as_uint(r0.x), hi = __float_as_uint(r0.y);
  uint64_t p = v.offsets[k];
  const uint64_t end = std::min<uint64_t>(p + count, v.pair_capacity);
  for (uint32_t ty = lo >> 16; ty <= (hi >> 16); ++ty) {
    for (uint32_t tx = lo & 0xffffu; tx <= (hi & 0xffffu) && p < end; ++tx, ++p) {
      v.emit_tiles[p] = ty * v.f.tiles_x + tx;
      v.emit_gaussians[p] = i;
    }
  }
}

__global__ void __launch_bounds__(256) k_view_composite(ViewArgs a) {
#pragma clang fp contract(off)
  __shared__ float4 s_a[kBatch], s_b[kBatch];
  __shared__ float2 s_c[kBatch];
  __shared__ uint64_t s_mask[4][kBatch / 64];
  const gsh::ViewDev& v = a.views[0];
  const gs_hip_frame& f = v.f;
  const uint32_t tile = blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), lane = int(threadIdx.x & 63u);
  const int tx0 = int(tile % f.tiles_x) * kTile, ty0 = int(tile / f.tiles_x) * kTile;
  const int px = tx0 + (wave & 1) * 8 + (lane & 7), py = ty0 + (wave >> 1) * 8 + (lane >> 3);
  const bool inside = px < int(f.cam.width) && py < int(f.cam.height);
  const float cx = float(px) + 0.5f, cy = float(py) + 0.5f;
  const uint2 range = v.ranges[tile];
  // The target holds 8-bit values: keep them as floats k in 0..255.
  float cr = rintf(fminf(fmaxf(a.clear[0], 0.0f), 1.0f) * 255.0f);
  float cg = rintf(fminf(fmaxf(a.clear[1], 0.0f), 1.0f) * 255.0f);
  float cb = rintf(fminf(fmaxf(a.clear[2], 0.0f), 1.0f) * 255.0f);
  float ca = rintf(fminf(fmaxf(a.clear[3], 0.0f), 1.0f) * 255.0f);
  for (uint32_t start = range.x; start < range.y; start += kBatch) {
    const uint32_t n = min(uint32_t(kBatch), range.y - start);
    uint32_t hits = 0u;
    if (threadIdx.x < n) {
      const uint32_t g = v.pair_gaussians[start + threadIdx.x];
      const float4* r = v.records + size_t(g) * 3u;
      const float4 ra = r[0], rb = r[1], rc = r[2];
      s_a[threadIdx.x] = ra;
      s_b[threadIdx.x] = rb;
      s_c[threadIdx.x] = make_float2(rc.x, rc.w);
      const uint32_t lo = __float_as_uint(rc.y), hi = __float_as_uint(rc.z);
      for (int q = 0; q < 4; ++q) {
        const int qx = tx0 + (q & 1) * 8, qy = ty0 + (q >> 1) * 8;
        hits |= bbox_hits(lo, hi, qx, qy, qx + 7, qy + 7) ? (1u << q) : 0u;
      }
    }
    for (int q = 0; q < 4; ++q) {
      const uint64_t m = __ballot((hits >> q) & 1u);
      if (lane == 0) s_mask[q][wave] = m;
    }
    __syncthreads();
    for (int chunk = 0; chunk < kBatch / 64; ++chunk) {
      uint64_t m = uniform_u64(s_mask[wave][chunk]);
      while (m != 0ull) {
        const int j = chunk * 64 + __builtin_ctzll(m);
        m &= m - 1ull;
        const float4 sa = s_a[j], sb = s_b[j];
        const float dx = cx - sa.x, dy = cy - sa.y;
        const float lx = dx * sa.z + dy * sa.w, ly = dx * sb.x + dy * sb.y;
        const float alpha = fminf(sb.z * gs_expf(-(lx * lx + ly * ly)), kMaxAlpha);
        if (fabsf(lx) <= 2.0f && fabsf(ly) <= 2.0f && alpha >= kMinAlpha) {
          const float2 sc = s_c[j];
          const float keep = 1.0f - alpha;
          cr = rintf(fminf(fminf(sb.w * alpha, 1.0f) * 255.0f + keep * cr, 255.0f));
          cg = rintf(fminf(fminf(sc.x * alpha, 1.0f) * 255.0f + keep * cg, 255.0f));
          cb = rintf(fminf(fminf(sc.y * alpha, 1.0f) * 255.0f + keep * cb, 255.0f));
          ca = rintf(fminf(alpha * 255.0f + keep * ca, 255.0f));
        }
      }
    }
    __syncthreads();
  }
  if (inside) {
    uint8_t* o = a.out + (size_t(py) * f.cam.width + px) * 4u;
    o[0] = uint8_t(cr), o[1] = uint8_t(cg), o[2] = uint8_t(cb), o[3] = uint8_t(ca);
  }
}

}  // namespace

extern "C" {

int gs_hip_view_render(gs_hip_workspace* ws, const gs_hip_frame* f, const float clear[4], uint32_t flags,
                       uint8_t* rgba8_out, uint32_t* visible) {
  if (ws == nullptr || f == nullptr || clear == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "view_render: null");
  if (f->slots > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "view_render: slots exceed capacity");
  if (f->cam.width != ws->width || f->cam.height != ws->height || f->tiles_x != ws->tiles_x ||
      f->tiles_y != ws->tiles_y) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "view_render: camera size differs from the workspace");
  }
  gsh::View& v = gsh::view0(ws);
  const size_t pixels = size_t(ws->width) * ws->height;
  ViewArgs a;
  a.params = reinterpret_cast<const float4*>(ws->params);
  a.sh = ws->sh;
  a.views = ws->dev_views;
  a.capacity = ws->capacity;
  a.sh_coeffs = ws->sh_coeffs;
  a.antialiased = (flags & GS_HIP_VIEW_ANTIALIASED) != 0u ? 1u : 0u;
  for (int c = 0; c < 4; ++c) a.clear[c] = clear[c];
  // The image goes to the view's image buffer (RGBA8 fits in its float4s), the count to totals[8].
  a.out = reinterpret_cast<uint8_t*>(v.image);
  a.visible = v.totals + 8;
  for (int attempt = 0;; ++attempt) {
    int rc = gsh::upload_table(ws, f, 1, nullptr, 0);  // the viewer bins by its 16x16 tiles
    if (rc == 0) rc = gsh::fill_u32(ws, a.visible, 1, 0u);
    if (rc != 0) return rc;
    if (ws->capacity > 0) {
      k_view_project<<<gsh::div_up(ws->capacity, 256), 256, 0, ws->stream>>>(a);
      GS_CHECK_LAUNCH("k_view_project");
    }
    rc = gsh::launch_depth_sort_and_scan(ws, 1, true);
    if (rc == 0) rc = gsh::launch_clear_ranges(ws, 1);
    if (rc != 0) return rc;
    if (ws->capacity > 0) {
      k_view_emit<<<gsh::div_up(ws->capacity, 256), 256, 0, ws->stream>>>(ws->dev_views);
      GS_CHECK_LAUNCH("k_view_emit");
    }
    rc = gsh::launch_tile_sort_and_ranges(ws, 1);
    if (rc != 0) return rc;
    k_view_composite<<<gsh::tile_count(ws), 256, 0, ws->stream>>>(a);
    GS_CHECK_LAUNCH("k_view_composite");
    GS_HIP_TRY(hipMemcpyAsync(ws->host_totals, v.totals, 64, hipMemcpyDeviceToHost, ws->stream));
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
    const uint64_t pairs = ws->host_totals[0];
    if (pairs <= v.pair_capacity) break;
    if (attempt > 4) return gsh::set_error(GS_HIP_E_CAPACITY, "view_render: pair buffers keep overflowing");
    ++ws->overflow_steps;
    rc = gsh::reserve_pairs(ws, pairs);
    if (rc != 0) return rc;
  }
  if (visible != nullptr) *visible = ws->host_totals[8];
  if (rgba8_out != nullptr) {
    GS_HIP_TRY(hipMemcpyAsync(rgba8_out, v.image, pixels * 4, hipMemcpyDeviceToHost, ws->stream));
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  }
  return 0;
}

}  // extern "C"
