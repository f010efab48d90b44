// EWA projection of every slot for every view of a step: screen mean,
// dilated 2D covariance and its conic, SH colour, opacity-aware footprint
// tile count, raster bbox and the depth sort key.  Follows
// engine/shaders/train_project.comp:26-112.
//
// One thread per slot handles all the step's views, so the 64 B of
// parameters and the 12 B per SH coefficient (coefficient-major planes: a
// wave reads 768 B contiguous per coefficient) are read once per step, not
// once per view.  Per view it writes a 48 B raster record, a 32 B footprint
// (all the pair emitter reads of a slot), tiles, misc, key and order, and
// clears the slot's nine GRAD2D accumulators.
// HBM-bound.
#include "gs_common.h"
#include "gs_internal.h"

namespace {

using namespace gsd;

struct ProjectArgs {
  const float4* __restrict__ params;
  const float* __restrict__ sh;
  const gsh::ViewDev* __restrict__ views;
  uint32_t capacity;
  uint32_t sh_coeffs;  // stored per slot
  uint32_t count;      // views
};

// What a slot's projection needs that no view changes: opacity, its
// footprint threshold, and M = R S (the normalised quaternion's rotation,
// columns scaled by exp(log scale)); computed once per step, not per view.
struct SlotShape {
  float opacity, thr;
  float m00, m01, m02, m10, m11, m12, m20, m21, m22;
};

__device__ __forceinline__ SlotShape slot_shape(float4 p0, float4 p1, float4 q) {
#pragma clang fp contract(off)
  SlotShape s;
  s.opacity = gs_sigmoid(p0.w);
  s.thr = footprint_threshold(s.opacity);
  const float qlen = fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
  const float w = q.x / qlen, x = q.y / qlen, y = q.z / qlen, z = q.w / qlen;
  const float sx = gs_expf(p1.x), sy = gs_expf(p1.y), sz = gs_expf(p1.z);
  s.m00 = (1.0f - 2.0f * (y * y + z * z)) * sx;
  s.m01 = (2.0f * (x * y - w * z)) * sy;
  s.m02 = (2.0f * (x * z + w * y)) * sz;
  s.m10 = (2.0f * (x * y + w * z)) * sx;
  s.m11 = (1.0f - 2.0f * (x * x + z * z)) * sy;
  s.m12 = (2.0f * (y * z - w * x)) * sz;
  s.m20 = (2.0f * (x * z - w * y)) * sx;
  s.m21 = (2.0f * (y * z + w * x)) * sy;
  s.m22 = (1.0f - 2.0f * (x * x + y * y)) * sz;
  return s;
}

// One view's projection of slot i (the arithmetic of the single-view
// kernel, term for term, with contraction off: records, tile counts and keys
// are bit-exact with the oracle's fp32 restatement).
__device__ __forceinline__ void project_view(const gsh::ViewDev& vd, const ProjectArgs& a, uint32_t i, float4 p0,
                                             float4 p1, const SlotShape& sh, float4 p3,
                                             const float (&rest)[GS_HIP_MAX_SH_REST][3], float log_depth_range) {
#pragma clang fp contract(off)
  const gs_hip_frame& f = vd.f;
  for (int c = 0; c < GS_HIP_GRAD2D_FLOATS; ++c) gmem(vd.grad2d)[size_t(i) * GS_HIP_GRAD2D_FLOATS + c] = 0.0f;
  const float* R = f.cam.rotation;
  // Camera-space position, left to right sums.
  const float tx = R[0] * p0.x + R[1] * p0.y + R[2] * p0.z + f.cam.translation[0];
  const float ty = R[3] * p0.x + R[4] * p0.y + R[5] * p0.z + f.cam.translation[1];
  const float tz = R[6] * p0.x + R[7] * p0.y + R[8] * p0.z + f.cam.translation[2];
  const float opacity = sh.opacity;

  uint32_t tiles = 0u;
  Record rec;
  rec.a = rec.b = rec.c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 foot = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (p1.w > 0.5f && tz > kNear && opacity >= kMinAlpha) {
    const float fx = f.cam.fx, fy = f.cam.fy, cx = f.cam.cx, cy = f.cam.cy;
    const float inv_z = 1.0f / tz;
    const float mx = fx * tx * inv_z + cx;
    const float my = fy * ty * inv_z + cy;
    // Jacobian of the perspective map, with the ray clamped to 1.3x the frustum.
    const float lim_x = vd.lim_x, lim_y = vd.lim_y;
    const float rx = fminf(fmaxf(tx * inv_z, -lim_x), lim_x);
    const float ry = fminf(fmaxf(ty * inv_z, -lim_y), lim_y);
    const float j00 = fx * inv_z, j02 = -fx * rx * inv_z;
    const float j11 = fy * inv_z, j12 = -fy * ry * inv_z;
    // T = J W (2x3), W the world -> camera rotation.
    const float t00 = j00 * R[0] + j02 * R[6], t01 = j00 * R[1] + j02 * R[7], t02 = j00 * R[2] + j02 * R[8];
    const float t10 = j11 * R[3] + j12 * R[6], t11 = j11 * R[4] + j12 * R[7], t12 = j11 * R[5] + j12 * R[8];
    const float m00 = sh.m00, m01 = sh.m01, m02 = sh.m02, m10 = sh.m10, m11 = sh.m11, m12 = sh.m12;
    const float m20 = sh.m20, m21 = sh.m21, m22 = sh.m22;
    // V = T M; cov2D = V V^T (= T M M^T T^T), dilated by 0.3 px^2.
    const float v00 = t00 * m00 + t01 * m10 + t02 * m20, v01 = t00 * m01 + t01 * m11 + t02 * m21,
                v02 = t00 * m02 + t01 * m12 + t02 * m22;
    const float v10 = t10 * m00 + t11 * m10 + t12 * m20, v11 = t10 * m01 + t11 * m11 + t12 * m21,
                v12 = t10 * m02 + t11 * m12 + t12 * m22;
    const float ca = v00 * v00 + v01 * v01 + v02 * v02 + kDilation;
    const float cb = v00 * v10 + v01 * v11 + v02 * v12;
    const float cc = v10 * v10 + v11 * v11 + v12 * v12 + kDilation;
    const float det = ca * cc - cb * cb;
    if (det > 0.0f) {
      const float mid = 0.5f * (ca + cc);
      const float lambda = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
      const float radius = ceilf(3.0f * sqrtf(lambda));
      const float thr = sh.thr;
      int row_first, row_last;
      footprint_rows(my, cc, thr, f.cam.height, row_first, row_last);
      if (row_first <= row_last) {
        const int side_log2 = 4 + int(vd.bin_shift);
        const RowSpan span = row_span(mx, my, ca, cb, cc, thr);
        for (uint32_t tyi = uint32_t(row_first) >> side_log2; tyi <= (uint32_t(row_last) >> side_log2); ++tyi) {
          uint32_t x0, x1;
          span_row_tiles(span, row_first, row_last, tyi, f.cam.width, x0, x1, side_log2);
          tiles += x1 - x0;
        }
      }
      if (tiles > 0u) {
        // SH colour in the direction from the camera centre.
        const float dx = p0.x - f.cam.position[0], dy = p0.y - f.cam.position[1], dz = p0.z - f.cam.position[2];
        const float dlen = sqrtf(dx * dx + dy * dy + dz * dz);
        float basis[16];
        sh_basis(dx / dlen, dy / dlen, dz / dlen, f.sh_degree, basis);
        float cr = basis[0] * p3.x, cg = basis[0] * p3.y, cbl = basis[0] * p3.z;
        const int coeffs = min(sh_rest_coeffs(f.sh_degree), int(a.sh_coeffs));
#pragma unroll
        for (int k = 0; k < GS_HIP_MAX_SH_REST; ++k) {  // unrolled: rest[] stays in registers
          if (k < coeffs) {
            cr = cr + basis[k + 1] * rest[k][0];
            cg = cg + basis[k + 1] * rest[k][1];
            cbl = cbl + basis[k + 1] * rest[k][2];
          }
        }
        cr = cr + 0.5f;
        cg = cg + 0.5f;
        cbl = cbl + 0.5f;
        const uint32_t clamped = (cr < 0.0f ? 1u : 0u) | (cg < 0.0f ? 2u : 0u) | (cbl < 0.0f ? 4u : 0u);
        const float inv_det = 1.0f / det;
        // Pixel bbox of the footprint ellipse: half extents sqrt(t a), sqrt(t c).
        int bx0, bx1, by0, by1;
        const float ex = sqrtf(thr * ca);
        footprint_pixels(mx - 0.5f - ex, mx - 0.5f + ex, f.cam.width, bx0, bx1);
        by0 = row_first;
        by1 = row_last;
        rec.a = make_float4(mx, my, cc * inv_det, -cb * inv_det);
        rec.b = make_float4(ca * inv_det, opacity, fmaxf(cr, 0.0f), fmaxf(cg, 0.0f));
        rec.c = make_float4(fmaxf(cbl, 0.0f), __uint_as_float(uint32_t(bx0) | (uint32_t(by0) << 16)),
                            __uint_as_float(uint32_t(bx1) | (uint32_t(by1) << 16)),
                            __uint_as_float((uint32_t(fminf(radius, 4.0e8f)) << 3) | clamped));
        foot = make_float4(ca, cb, cc, thr);
      }
    }
  }
  float4* r = vd.records + size_t(i) * 3u;
  gstore(r, rec.a);
  gstore(r + 1, rec.b);
  gstore(r + 2, rec.c);
  gstore(vd.footprint + size_t(i) * 2u, foot);
  gstore(vd.footprint + size_t(i) * 2u + 1u, make_float4(rec.a.x, rec.a.y, __uint_as_float(tiles), 0.0f));
  gmem(vd.tiles)[i] = tiles;
  gmem(vd.misc)[i] = __float_as_uint(rec.c.w);
  const bool depth16 = (f.flags & GS_HIP_FLAG_DEPTH16) != 0u;
  uint32_t key = depth16 ? 0xffffu : 0xffffffffu;
  if (tiles > 0u) {
    if (depth16) {
      // The reference's 16-bit log depth (train_project.comp:110) in the lower half; the sort
      // ranks those 16 bits only, and the tile count (< 2^16) rides in the upper half.
      const float u = gs_logf(tz / kNear) / log_depth_range;
      key = uint32_t(fminf(fmaxf(u, 0.0f), 1.0f) * 65535.0f) | (tiles << 16);
    } else {
      key = __float_as_uint(tz);  // z > 0: the bits order like the values
    }
  }
  gmem(vd.depth_keys)[i] = key;
  gmem(vd.depth_order)[i] = i;
}

// One thread per slot of scene blockIdx.y (its model and views from the tables).
__global__ void __launch_bounds__(256) k_project(const gsh::SceneDev* __restrict__ scenes,
                                                 const gsh::ViewDev* __restrict__ views) {
  const gsh::SceneDev& sc = scenes[blockIdx.y];
  ProjectArgs a;
  a.params = sc.params;
  a.sh = sc.sh;
  a.views = views + sc.first_view;
  a.capacity = sc.capacity;
  a.sh_coeffs = sc.sh_coeffs;
  a.count = sc.views;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.views[0].f.slots) return;
  const float4 p0 = gload(a.params + i * 4u), p1 = gload(a.params + i * 4u + 1u);
  const float4 q = gload(a.params + i * 4u + 2u), p3 = gload(a.params + i * 4u + 3u);
  // SH rest of a live slot, once for all views.
  float rest[GS_HIP_MAX_SH_REST][3];
#pragma unroll
  for (int k = 0; k < GS_HIP_MAX_SH_REST; ++k) rest[k][0] = rest[k][1] = rest[k][2] = 0.0f;
  if (p1.w > 0.5f) {
#pragma unroll
    for (int k = 0; k < GS_HIP_MAX_SH_REST; ++k) {
      if (k < int(a.sh_coeffs)) {
        const GS_GLOBAL float* s = gmem(a.sh) + (size_t(k) * a.capacity + i) * 3u;
        rest[k][0] = s[0];
        rest[k][1] = s[1];
        rest[k][2] = s[2];
      }
    }
  }
  const SlotShape shape = slot_shape(p0, p1, q);
  const float log_depth_range = gs_logf(kFar / kNear);
  for (uint32_t v = 0; v < a.count; ++v) project_view(a.views[v], a, i, p0, p1, shape, p3, rest, log_depth_range);
}

}  // namespace

namespace gsh {
int launch_project(gs_hip_workspace* ws, uint32_t count) {
  (void)count;
  if (ws->shape.max_capacity == 0) return 0;
  timer_begin(ws, "project");
  k_project<<<dim3(div_up(ws->shape.max_capacity, 256), ws->shape.scenes), 256, 0, ws->stream>>>(ws->dev_scenes,
                                                                                                  ws->dev_views);
  GS_CHECK_LAUNCH("k_project");
  timer_end(ws);
  return 0;
}
}  // namespace gsh
