/*
 * CPU oracle for the 3DGS training step — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline); the
 * product path never calls it.
 *
 * Two restatements of the reference engine's algorithm:
 *
 *  fp64: the double-precision reference of engine/src/train_selftest.cpp:38-283
 *        (projection, brute-force compositing over every visible gaussian,
 *        separable zero-padded SSIM, the loss) plus an analytic backward
 *        (per-pixel back-to-front adjoint of engine/shaders/train_common.glsl:
 *        215-253, then the projection adjoint of train_step.glsl:72-220),
 *        pinned by central finite differences of the fp64 loss.
 *
 *  fp32: the projection chain in single precision in a fixed evaluation
 *        order with contraction off (built with -ffp-contract=off), using
 *        exp/log built from correctly rounded operations (Cody-Waite
 *        reduction + Taylor series for exp, the atanh series for log).  The
 *        kernels evaluate the same sequence, so records, tile counts, depth
 *        keys, pairs and ranges match bit for bit.  Binning is the reference's
 *        order: pairs sorted by (tile, depth key, gaussian index)
 *        (engine/tests/test_trainer.cpp:197-233).  The tile compositor
 *        (train_forward.comp:77-117) runs on those records with libm expf.
 *
 * Parity status per function (what pins it, checked in tests/test_oracle.py):
 *   orc_expf, orc_logf        libm exp/log to <= 2 ulp over sweeps
 *   orc_project64             closed forms: an isotropic splat on the axis (the
 *                             reference's single-splat falloff KAT), and rotated,
 *                             anisotropic splats head-on and off-axis with the
 *                             Jacobian's 1.3x clamp active against a first-
 *                             principles numpy construction (axis-angle rotation,
 *                             J V R S^2 R^T V^T J^T + 0.3 I) that a transposed
 *                             rotation or a missing clamp demonstrably misses
 *   orc_project32             orc_project64 to 1e-5 relative
 *   orc_sh_basis64            scipy's complex spherical harmonics (real form)
 *   orc_footprint_tiles32     brute-force per-pixel alpha >= 1/255 rule
 *                             (test_trainer.cpp:235-289)
 *   orc_footprint_bins32      the same rule at 32 and 64 px bins, and each bin's
 *                             set holds the union of its 16 px tiles' sets
 *   orc_bin32, orc_bin32_bins Python sorted() on (bin, key, index) tuples
 *   orc_bin_counts32          orc_footprint_bins32 slot by slot
 *   orc_render64              closed forms (one splat, two splats in depth order);
 *                             its tile lists equal orc_render64_brute bit for bit
 *   orc_forward32             orc_render64 within fp32 tolerance
 *   orc_ssim64                an independent numpy/scipy.ndimage SSIM;
 *                             SSIM(x, x) = 1
 *   orc_loss_image64          central FD of itself (its gradient output)
 *   orc_backward32            orc_backward64 on small scenes (same image)
 *   orc_adjoint64             central FD of orc_loss64 through orc_backward64
 *   orc_backward64            central FD of orc_loss64 (h = 1e-4 .. 1e-6), and
 *                             orc_fd_frozen64 on every parameter of a 120-splat scene
 *   orc_fd_frozen64           orc_backward64 (tight), and plain central FD of
 *                             orc_loss64 wherever no 1/255 contour crosses a pixel
 *   orc_adam                  closed form; lr 0 leaves m/(1-b1) = g
 *   orc_densify_mark          the predicate of train_densify_mark.comp:28-53
 *   orc_half                  exhaustive: every fp16 value and the midpoints between
 *                             neighbours (ties to even)
 *   orc_view_render32         the reference viewer tests' closed forms (single-splat
 *                             falloff, depth order, axis-flip equivalence)
 *   orc_pose_adjoint64        central FD of orc_loss64 under camera shifts and turns
 *                             (SH degree 0), and the reference's identity with the
 *                             per-gaussian gradients (test_trainer.cpp:404-464)
 *
 * Multithreaded with OpenMP (the GPU box's host cores); every sum whose order
 * matters is taken in a fixed order, so results do not depend on the thread
 * count except for orc_backward32's per-gaussian sums, which add per-pair
 * partial sums in pair order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TILE 16
#define NEAR 0.02
#define FAR 1000.0
#define DILATION 0.3
#define MIN_ALPHA (1.0 / 255.0)
#define MAX_ALPHA 0.99
#define MIN_T 1e-4

/* Camera: rotation[9] row-major (scene -> camera, OpenCV axes), translation[3],
 * position[3], fx, fy, cx, cy, width, height — as gs::Camera. */
typedef struct {
  float rotation[9];
  float translation[3];
  float position[3];
  float fx, fy, cx, cy;
  uint32_t width, height;
} orc_camera;

static float i2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t f2i(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* ===================== fp32 exp / log (correctly rounded ops) ============== */
float orc_expf(float x) {
  if (x > 88.72f) return i2f(0x7f800000u);
  if (x < -103.97f) return 0.0f;
  const float n = rintf(x * 1.44269504f);
  float r = x - n * 0.693145751953125f;
  r = r - n * 1.428606765330187e-06f;
  float p = 1.98412698e-4f;
  p = p * r + 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  const int k = (int)n;
  const int k1 = k / 2, k2 = k - k1;
  p = p * i2f((uint32_t)(k1 + 127) << 23);
  return p * i2f((uint32_t)(k2 + 127) << 23);
}

float orc_logf(float x) {
  const uint32_t bits = f2i(x);
  int e = (int)((bits >> 23) & 0xffu) - 127;
  float m = i2f((bits & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = 9.09090909e-2f;
  p = p * s2 + 1.11111111e-1f;
  p = p * s2 + 1.42857143e-1f;
  p = p * s2 + 2.0e-1f;
  p = p * s2 + 3.33333333e-1f;
  p = p * s2 + 1.0f;
  const float lm = (2.0f * s) * p;
  const float fe = (float)e;
  return fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
}

/* ===================== SH basis ============================================ */
/* Real SH to degree 3 in the 3DGS convention (train_common.glsl:84-109):
 * basis k is its normalisation times a polynomial in the direction, band by
 * band (m = -l..l within degree l). */
void orc_sh_basis64(const double d[3], int degree, double b[16]) {
  static const double norm0 = 0.28209479177387814, norm1 = 0.4886025119029199;
  static const double norm2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                                  -1.0925484305920792, 0.5462742152960396};
  static const double norm3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                                  -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
  const double x = d[0], y = d[1], z = d[2];
  const double xx = x * x, yy = y * y, zz = z * z;
  const double band1[3] = {-y, z, -x};
  const double band2[5] = {x * y, y * z, 2 * zz - xx - yy, x * z, xx - yy};
  const double band3[7] = {y * (3 * xx - yy),     x * y * z,         y * (4 * zz - xx - yy), z * (2 * zz - 3 * xx - 3 * yy),
                           x * (4 * zz - xx - yy), z * (xx - yy), x * (xx - 3 * yy)};
  memset(b, 0, sizeof(double) * 16);
  b[0] = norm0;
  for (int k = 0; degree >= 1 && k < 3; ++k) b[1 + k] = norm1 * band1[k];
  for (int k = 0; degree >= 2 && k < 5; ++k) b[4 + k] = norm2[k] * band2[k];
  for (int k = 0; degree >= 3 && k < 7; ++k) b[9 + k] = norm3[k] * band3[k];
}

static void sh_basis32(float x, float y, float z, uint32_t degree, float b[16]) {
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

static int rest_coeffs(int degree) { return degree <= 0 ? 0 : (degree + 1) * (degree + 1) - 1; }

/* ===================== fp32 footprint (train_common.glsl:163-213) ========== */
static void footprint_pixels(float lo, float hi, uint32_t size, int* first, int* last) {
  const int n = (int)size;
  const float fn = (float)n;
  int a = (int)fminf(fmaxf(ceilf(lo), -1.0f), fn);
  int b = (int)fminf(fmaxf(floorf(hi), -1.0f), fn);
  *first = a < 0 ? 0 : (a > n ? n : a);
  *last = b < -1 ? -1 : (b > n - 1 ? n - 1 : b);
}

static void footprint_rows(float my, float c, float t, uint32_t height, int* first, int* last) {
  const float ey = sqrtf(t * c);
  const float v = my - 0.5f;
  footprint_pixels(v - ey, v + ey, height, first, last);
}

/* Bins of `bin` pixels (TILE for the reference's tiles) the ellipse reaches
 * in bin row ty. */
static void footprint_row_tiles(float mx, float my, float a, float b, float c, float t, int row_first, int row_last,
                                uint32_t ty, uint32_t width, uint32_t bin, uint32_t* x0, uint32_t* x1) {
  const float u = mx - 0.5f, v = my - 0.5f;
  const int top = (int)(ty * bin), bottom = (int)(ty * bin + bin) - 1;
  const float y0 = (float)(top > row_first ? top : row_first) - v;
  const float y1 = (float)(bottom < row_last ? bottom : row_last) - v;
  const float k = b / c;
  const float w = (a * c - b * b) / c;
  const float side = b * sqrtf(t / a);
  const float dr = fminf(fmaxf(side, y0), y1);
  const float dl = fminf(fmaxf(-side, y0), y1);
  const float x_max = k * dr + sqrtf(fmaxf(w * (t - dr * dr / c), 0.0f));
  const float x_min = k * dl - sqrtf(fmaxf(w * (t - dl * dl / c), 0.0f));
  int first, last;
  footprint_pixels(u + x_min, u + x_max, width, &first, &last);
  if (first > last) {
    *x0 = *x1 = 0u;
    return;
  }
  *x0 = (uint32_t)first / bin;
  *x1 = (uint32_t)last / bin + 1u;
}

/* Bins of `side` pixels (by * bins_x + bx, row by row) the footprint reaches;
 * returns the count, writes at most max_out of them.  side = TILE gives the
 * reference's tiles; the kernels' coarser pair bins (32, 64) use the same
 * rule per bin row. */
uint32_t orc_footprint_bins32(float mx, float my, float a, float b, float c, float t, uint32_t width,
                              uint32_t height, uint32_t side, uint32_t* out, uint32_t max_out) {
  const uint32_t bins_x = (width + side - 1) / side;
  int rf, rl;
  footprint_rows(my, c, t, height, &rf, &rl);
  uint32_t n = 0;
  if (rf > rl) return 0;
  for (uint32_t ty = (uint32_t)rf / side; ty <= (uint32_t)rl / side; ++ty) {
    uint32_t x0, x1;
    footprint_row_tiles(mx, my, a, b, c, t, rf, rl, ty, width, side, &x0, &x1);
    for (uint32_t x = x0; x < x1; ++x) {
      if (out != NULL && n < max_out) out[n] = ty * bins_x + x;
      ++n;
    }
  }
  return n;
}

uint32_t orc_footprint_tiles32(float mx, float my, float a, float b, float c, float t, uint32_t width,
                               uint32_t height, uint32_t* out, uint32_t max_out) {
  return orc_footprint_bins32(mx, my, a, b, c, t, width, height, TILE, out, max_out);
}

/* ===================== fp32 projection (train_project.comp:26-112) ========= */
/* params [n][16]; sh [n][coeffs_in][3]; outputs records [n][12], footprint
 * [n][4], tiles [n], keys [n].  Same order of operations as the kernel. */
void orc_project32(const float* params, const float* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                   uint32_t sh_degree, int depth16, float* records, float* footprint, uint32_t* tiles,
                   uint32_t* keys) {
  const float* R = cam->rotation;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = params + (size_t)i * 16;
    float* rec = records + (size_t)i * 12;
    float* foot = footprint + (size_t)i * 4;
    memset(rec, 0, 48);
    memset(foot, 0, 16);
    const float tx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + cam->translation[0];
    const float ty = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + cam->translation[1];
    const float tz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + cam->translation[2];
    const float opacity = 1.0f / (1.0f + orc_expf(-p[3]));
    uint32_t count = 0;
    if (p[7] > 0.5f && tz > 0.02f && opacity >= 1.0f / 255.0f) {
      const float fx = cam->fx, fy = cam->fy, cx = cam->cx, cy = cam->cy;
      const float W = (float)cam->width, H = (float)cam->height;
      const float inv_z = 1.0f / tz;
      const float mx = fx * tx * inv_z + cx;
      const float my = fy * ty * inv_z + cy;
      const float lim_x = 1.3f * fmaxf(cx, W - cx) / fx;
      const float lim_y = 1.3f * fmaxf(cy, H - cy) / fy;
      const float rx = fminf(fmaxf(tx * inv_z, -lim_x), lim_x);
      const float ry = fminf(fmaxf(ty * inv_z, -lim_y), lim_y);
      const float j00 = fx * inv_z, j02 = -fx * rx * inv_z;
      const float j11 = fy * inv_z, j12 = -fy * ry * inv_z;
      const float t00 = j00 * R[0] + j02 * R[6], t01 = j00 * R[1] + j02 * R[7], t02 = j00 * R[2] + j02 * R[8];
      const float t10 = j11 * R[3] + j12 * R[6], t11 = j11 * R[4] + j12 * R[7], t12 = j11 * R[5] + j12 * R[8];
      const float* q = p + 8;
      const float qlen = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), 1e-12f);
      const float w = q[0] / qlen, x = q[1] / qlen, y = q[2] / qlen, z = q[3] / qlen;
      const float sx = orc_expf(p[4]), sy = orc_expf(p[5]), sz = orc_expf(p[6]);
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
      const float ca = v00 * v00 + v01 * v01 + v02 * v02 + 0.3f;
      const float cb = v00 * v10 + v01 * v11 + v02 * v12;
      const float cc = v10 * v10 + v11 * v11 + v12 * v12 + 0.3f;
      const float det = ca * cc - cb * cb;
      if (det > 0.0f) {
        const float mid = 0.5f * (ca + cc);
        const float lambda = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
        const float radius = ceilf(3.0f * sqrtf(lambda));
        const float t1 = 2.0f * orc_logf(opacity * 255.0f);
        const float thr = t1 * 1.0001f + 1e-4f;
        int rf, rl;
        footprint_rows(my, cc, thr, cam->height, &rf, &rl);
        if (rf <= rl) {
          for (uint32_t tyi = (uint32_t)rf / TILE; tyi <= (uint32_t)rl / TILE; ++tyi) {
            uint32_t x0, x1;
            footprint_row_tiles(mx, my, ca, cb, cc, thr, rf, rl, tyi, cam->width, TILE, &x0, &x1);
            count += x1 - x0;
          }
        }
        if (count > 0u) {
          const float dx = p[0] - cam->position[0], dy = p[1] - cam->position[1], dz = p[2] - cam->position[2];
          const float dlen = sqrtf(dx * dx + dy * dy + dz * dz);
          float basis[16];
          sh_basis32(dx / dlen, dy / dlen, dz / dlen, sh_degree, basis);
          float cr = basis[0] * p[12], cg = basis[0] * p[13], cbl = basis[0] * p[14];
          int coeffs = rest_coeffs((int)sh_degree);
          if (coeffs > (int)coeffs_in) coeffs = (int)coeffs_in;
          for (int k = 0; k < coeffs; ++k) {
            const float* s = sh + ((size_t)i * coeffs_in + (size_t)k) * 3;
            cr = cr + basis[k + 1] * s[0];
            cg = cg + basis[k + 1] * s[1];
            cbl = cbl + basis[k + 1] * s[2];
          }
          cr = cr + 0.5f;
          cg = cg + 0.5f;
          cbl = cbl + 0.5f;
          const uint32_t clamped = (cr < 0.0f ? 1u : 0u) | (cg < 0.0f ? 2u : 0u) | (cbl < 0.0f ? 4u : 0u);
          const float inv_det = 1.0f / det;
          int bx0, bx1;
          const float ex = sqrtf(thr * ca);
          footprint_pixels(mx - 0.5f - ex, mx - 0.5f + ex, cam->width, &bx0, &bx1);
          rec[0] = mx;
          rec[1] = my;
          rec[2] = cc * inv_det;
          rec[3] = -cb * inv_det;
          rec[4] = ca * inv_det;
          rec[5] = opacity;
          rec[6] = fmaxf(cr, 0.0f);
          rec[7] = fmaxf(cg, 0.0f);
          rec[8] = fmaxf(cbl, 0.0f);
          rec[9] = i2f((uint32_t)bx0 | ((uint32_t)rf << 16));
          rec[10] = i2f((uint32_t)bx1 | ((uint32_t)rl << 16));
          rec[11] = i2f(((uint32_t)fminf(radius, 4.0e8f) << 3) | clamped);
          foot[0] = ca;
          foot[1] = cb;
          foot[2] = cc;
          foot[3] = thr;
        }
      }
    }
    tiles[i] = count;
    uint32_t key = 0xffffffffu;
    if (count > 0u) {
      if (depth16) {
        const float u = orc_logf(tz / 0.02f) / orc_logf(1000.0f / 0.02f);
        key = (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * 65535.0f);
      } else {
        key = f2i(tz);
      }
    }
    keys[i] = key;
  }
}

/* ===================== binning (pairs by (tile, key, index)) ============== */
typedef struct {
  uint32_t tile, key, gaussian;
} pair_t;

/* Within one tile: by depth key, then gaussian index. */
static int pair_cmp(const void* a, const void* b) {
  const pair_t* x = (const pair_t*)a;
  const pair_t* y = (const pair_t*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  if (x->gaussian != y->gaussian) return x->gaussian < y->gaussian ? -1 : 1;
  return 0;
}

/* Emits every (tile, gaussian) pair from records/footprints, sorts them by
 * (tile, key, index) and fills ranges [(tiles+1)][2].  Returns the pair
 * count; writes up to max_pairs pairs.  (A slot whose footprint walk finds
 * fewer tiles than its count pads with the overflow tile `tiles`.) */
uint64_t orc_bin32_bins(const float* records, const float* footprint, const uint32_t* tiles, const uint32_t* keys,
                        uint32_t n, uint32_t width, uint32_t height, uint32_t side, uint32_t* pair_tiles,
                        uint32_t* pair_gaussians, uint64_t max_pairs, uint32_t* ranges) {
  const uint32_t tiles_x = (width + side - 1) / side, tiles_y = (height + side - 1) / side;
  const uint32_t T = tiles_x * tiles_y;
  uint64_t* first = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  first[0] = 0;
  for (uint32_t i = 0; i < n; ++i) first[i + 1] = first[i] + tiles[i];
  const uint64_t total = first[n];
  pair_t* emitted = (pair_t*)malloc(sizeof(pair_t) * (total > 0 ? total : 1));
  pair_t* pairs = (pair_t*)malloc(sizeof(pair_t) * (total > 0 ? total : 1));
#pragma omp parallel
  {
    uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * (T + 1));
#pragma omp for schedule(dynamic, 1024)
    for (uint32_t i = 0; i < n; ++i) {
      if (tiles[i] == 0) continue;
      const float* r = records + (size_t)i * 12;
      const float* f = footprint + (size_t)i * 4;
      const uint32_t c = orc_footprint_bins32(r[0], r[1], f[0], f[1], f[2], f[3], width, height, side, buf, T + 1);
      for (uint32_t k = 0; k < tiles[i]; ++k) {
        pair_t* q = &emitted[first[i] + k];
        q->tile = k < c ? buf[k] : T;
        q->key = keys[i];
        q->gaussian = i;
      }
    }
    free(buf);
  }
  /* Stable counting sort by tile (pairs stay in gaussian order within a
   * tile), then each tile by (key, gaussian). */
  uint64_t* start = (uint64_t*)calloc((size_t)T + 2, sizeof(uint64_t));
  for (uint64_t s = 0; s < total; ++s) ++start[emitted[s].tile + 1];
  for (uint32_t t = 0; t <= T; ++t) start[t + 1] += start[t];
  uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)T + 1));
  memcpy(fill, start, sizeof(uint64_t) * ((size_t)T + 1));
  for (uint64_t s = 0; s < total; ++s) pairs[fill[emitted[s].tile]++] = emitted[s];
#pragma omp parallel for schedule(dynamic, 16)
  for (uint32_t t = 0; t <= T; ++t) {
    if (start[t + 1] - start[t] > 1) qsort(pairs + start[t], (size_t)(start[t + 1] - start[t]), sizeof(pair_t), pair_cmp);
  }
  memset(ranges, 0, sizeof(uint32_t) * 2 * (T + 1));
  for (uint32_t t = 0; t <= T; ++t) {
    if (start[t + 1] > start[t]) {
      ranges[2 * t] = (uint32_t)start[t];
      ranges[2 * t + 1] = (uint32_t)start[t + 1];
    }
  }
  const uint64_t kept = total < max_pairs ? total : max_pairs;
#pragma omp parallel for schedule(static)
  for (uint64_t s = 0; s < kept; ++s) {
    pair_tiles[s] = pairs[s].tile;
    pair_gaussians[s] = pairs[s].gaussian;
  }
  free(first);
  free(emitted);
  free(pairs);
  free(start);
  free(fill);
  return total;
}

uint64_t orc_bin32(const float* records, const float* footprint, const uint32_t* tiles, const uint32_t* keys,
                   uint32_t n, uint32_t width, uint32_t height, uint32_t* pair_tiles, uint32_t* pair_gaussians,
                   uint64_t max_pairs, uint32_t* ranges) {
  return orc_bin32_bins(records, footprint, tiles, keys, n, width, height, TILE, pair_tiles, pair_gaussians,
                        max_pairs, ranges);
}

/* Bins of `side` pixels each slot's footprint reaches (0 for a slot the
 * projection culled: tiles[i] == 0). */
void orc_bin_counts32(const float* records, const float* footprint, const uint32_t* tiles, uint32_t n,
                      uint32_t width, uint32_t height, uint32_t side, uint32_t* counts) {
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = records + (size_t)i * 12;
    const float* f = footprint + (size_t)i * 4;
    counts[i] = tiles[i] == 0 ? 0u : orc_footprint_bins32(r[0], r[1], f[0], f[1], f[2], f[3], width, height, side,
                                                          NULL, 0);
  }
}

/* ===================== fp32 tile compositor (train_forward.comp:77-117) ===== */
/* A splat evaluation whose skip / clamp / stop decision sits near its
 * threshold (power > 0, alpha < 1/255, raw alpha >= 0.99): another fp32
 * evaluation order or exp (the GPU's, the Vulkan driver's) may decide it the
 * other way, so a comparison at the bit level is meaningless there.  "Near":
 * within 1e-5 relative in alpha, plus the rounding of the power, which is a
 * sum of terms that cancel (a dx^2 / 2, b dx dy, c dy^2 / 2: for a needle
 * splat far along its axis each is ~100 while the power is ~-5), so another
 * order of operations moves it by some ulp of the largest term: 2^-19 (16
 * ulp) of `mag`, their magnitudes' sum.  Pixels that evaluate such a splat
 * are reported as ambiguous. */
static int near_threshold(float power, float raw, float mag) {
  const float a = 1.0f / 255.0f, band = 1e-5f + 0x1p-19f * mag;
  return power > -1e-6f - 0x1p-19f * mag || fabsf(raw - a) <= band * a || fabsf(raw - 0.99f) <= band * 0.99f;
}

static float power_magnitude(const float* r, float dx, float dy) {
  return 0.5f * (fabsf(r[2]) * dx * dx + fabsf(r[4]) * dy * dy) + fabsf(r[3] * dx * dy);
}

/* Per pair, the power below which its splat's alpha is under 1/255:
 * ln(1 / (255 o)).  A walk skips a splat whose power is below that by more
 * than twice near_threshold's band without evaluating exp (it would be
 * skipped anyway, and no other evaluation could draw it). */
static float* faint_powers(const float* records, const uint32_t* pair_gaussians, uint64_t pairs) {
  float* cut = (float*)malloc(sizeof(float) * (pairs > 0 ? pairs : 1));
#pragma omp parallel for schedule(static)
  for (uint64_t k = 0; k < pairs; ++k) cut[k] = logf(1.0f / (255.0f * records[(size_t)pair_gaussians[k] * 12 + 5]));
  return cut;
}

static int clearly_faint(float power, float cut, float mag) { return power < cut - (2e-5f + 0x1p-18f * mag); }

/* image [H*W][4] = (C rgb, final T); last [H*W] = 1-based position of the
 * last composited splat in its tile's range; spread [H*W] (may be null): how
 * far another fp32 evaluation of the same splats can move the pixel's values
 * (each contributor's alpha uncertain by 1e-5 plus 16 ulp of its power's
 * terms, through its weight in C and T) -- ill-conditioned where needle
 * splats far along their axis cancel terms of ~1e5; ambiguous [H*W] (may be null):
 * 1 where the pixel met a splat at a decision threshold (near_threshold) or
 * stopped within 1e-5 relative of T = 1e-4. */
void orc_forward32(const float* records, const uint32_t* pair_gaussians, const uint32_t* ranges, uint32_t width,
                   uint32_t height, float* image, uint32_t* last, uint8_t* ambiguous, float* spread) {
  const uint32_t tiles_x = (width + TILE - 1) / TILE, tiles_y = (height + TILE - 1) / TILE;
  const uint32_t T = tiles_x * tiles_y;
  uint64_t total = 0, longest = 0;
  for (uint32_t t = 0; t <= T; ++t) {
    if (ranges[2 * t + 1] > total) total = ranges[2 * t + 1];
    if (t < T && ranges[2 * t + 1] - ranges[2 * t] > longest) longest = ranges[2 * t + 1] - ranges[2 * t];
  }
  float* cut = faint_powers(records, pair_gaussians, total);
#pragma omp parallel
  {
    /* A tile's records copied together (the same values, so the same
     * results; the pixels of a tile then read them from one block). */
    float* staged = (float*)malloc(sizeof(float) * 12 * (longest > 0 ? longest : 1));
#pragma omp for schedule(dynamic, 1)
    for (uint32_t tile = 0; tile < T; ++tile) {
      const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
      for (uint32_t k = start; k < end; ++k)
        memcpy(staged + (size_t)(k - start) * 12, records + (size_t)pair_gaussians[k] * 12, 48);
      const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
      for (uint32_t py = ty * TILE; py < (ty + 1) * TILE && py < height; ++py) {
        for (uint32_t px = tx * TILE; px < (tx + 1) * TILE && px < width; ++px) {
          const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
          float Tr = 1.0f, C[3] = {0.0f, 0.0f, 0.0f};
          uint32_t lst = 0;
          int amb = 0;
          double sp = 0.0;
          for (uint32_t k = start; k < end; ++k) {
            const float* r = staged + (size_t)(k - start) * 12;
            const float dx = r[0] - cx, dy = r[1] - cy;
            const float power = -0.5f * (r[2] * dx * dx + r[4] * dy * dy) - r[3] * dx * dy;
            const float mag = power_magnitude(r, dx, dy);
            if (clearly_faint(power, cut[k], mag)) continue;
            const float raw = r[5] * expf(power);
            amb |= near_threshold(power, raw, mag);
            if (power > 0.0f) continue;
            const float alpha = fminf(0.99f, raw);
            if (alpha < 1.0f / 255.0f) continue;
            const float next_T = Tr * (1.0f - alpha);
            amb |= fabsf(next_T - 1e-4f) <= 1e-5f * 1e-4f;
            if (next_T < 1e-4f) break;
            for (int c = 0; c < 3; ++c) C[c] += r[6 + c] * alpha * Tr;
            /* alpha's relative uncertainty under another evaluation (16 ulp of
             * the power's terms, 1e-5 for exp), times how far it moves C and T */
            const double cmax = fmax(fmax(r[6], r[7]), fmax(r[8], 1.0f));
            sp += (1e-5 + 0x1p-19 * (double)mag) * alpha * (Tr * 2.0 * cmax + 1.0 / (1.0 - alpha));
            Tr = next_T;
            lst = k - start + 1;
          }
          const size_t pix = (size_t)py * width + px;
          image[pix * 4 + 0] = C[0];
          image[pix * 4 + 1] = C[1];
          image[pix * 4 + 2] = C[2];
          image[pix * 4 + 3] = Tr;
          last[pix] = lst;
          if (ambiguous) ambiguous[pix] = (uint8_t)amb;
          if (spread) spread[pix] = (float)sp;
        }
      }
    }
    free(staged);
  }
  free(cut);
}

/* ===================== fp32 tile backward (train_backward_serial.glsl) ====== */
/* Per pixel, back to front over its tile's list from its last contributor,
 * with the forward's final T (image[..][3]) and dL/dC (dldc [H*W][4]); the
 * per-pair terms of splat_pixel_grad (train_common.glsl:228-253), summed per
 * pair over the tile's pixels in double and then per gaussian (in pair
 * order) into grad2d [n][9].
 * A pixel is ambiguous when its walk meets a splat at a decision threshold
 * (near_threshold): another fp32 evaluation may skip or clamp that splat the
 * other way and so shift every term of that pixel.  tainted [n] (may be null):
 * 1 for every gaussian with a term at an ambiguous pixel.  share [n][9] (may be
 * null), per gaussian and value, how far another fp32 evaluation can move the
 * sums: the |terms| at ambiguous pixels, plus every |term| times the pixel's
 * conditioning and the accumulation's rounding.  The walk recovers T by
 * dividing by 1 - alpha, so an alpha that another exp / evaluation order
 * makes differ by d (relative) moves the recovered T, and every term behind
 * it, by d alpha / (1 - alpha) relative; with d = 2^-20 (8 ulp) the pixel's
 * conditioning is 2^-20 times the sum of alpha / (1 - alpha) over its
 * contributors.  A sum of fp32 terms of either sign (the GPU's wave, LDS and
 * global float additions) rounds to a few ulp of the sum of their
 * magnitudes, which can far exceed the sum itself (a wide splat's conic
 * terms cancel): 2^-18 (32 ulp) of every |term|. */
void orc_backward32(const float* records, const uint32_t* pair_gaussians, const uint32_t* ranges, const float* image,
                    const uint32_t* last, const float* dldc, const float bg[3], uint32_t n, uint32_t width,
                    uint32_t height, double* grad2d, uint8_t* tainted, double* share) {
  const uint32_t tiles_x = (width + TILE - 1) / TILE, tiles_y = (height + TILE - 1) / TILE;
  const uint32_t T = tiles_x * tiles_y;
  const uint64_t pairs = ranges[2 * T + 1] > 0 ? ranges[2 * T + 1] : 0;
  uint64_t total = 0;
  for (uint32_t t = 0; t <= T; ++t)
    if (ranges[2 * t + 1] > total) total = ranges[2 * t + 1];
  (void)pairs;
  double* terms = (double*)calloc((size_t)(total > 0 ? total : 1) * 9, sizeof(double));
  double* amb_terms = share ? (double*)calloc((size_t)(total > 0 ? total : 1) * 9, sizeof(double)) : NULL;
  uint8_t* amb_pair = tainted ? (uint8_t*)calloc((size_t)(total > 0 ? total : 1), 1) : NULL;
  float* cut = faint_powers(records, pair_gaussians, total);
  uint64_t longest = 0;
  for (uint32_t t = 0; t < T; ++t)
    if (ranges[2 * t + 1] - ranges[2 * t] > longest) longest = ranges[2 * t + 1] - ranges[2 * t];
#pragma omp parallel
  {
  float* staged = (float*)malloc(sizeof(float) * 12 * (longest > 0 ? longest : 1));
#pragma omp for schedule(dynamic, 1)
  for (uint32_t tile = 0; tile < T; ++tile) {
    const uint32_t start = ranges[2 * tile];
    for (uint32_t k = start; k < ranges[2 * tile + 1]; ++k)
      memcpy(staged + (size_t)(k - start) * 12, records + (size_t)pair_gaussians[k] * 12, 48);
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    for (uint32_t py = ty * TILE; py < (ty + 1) * TILE && py < height; ++py) {
      for (uint32_t px = tx * TILE; px < (tx + 1) * TILE && px < width; ++px) {
        const size_t pix = (size_t)py * width + px;
        const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
        const float T_final = image[pix * 4 + 3];
        const float* dC = dldc + pix * 4;
        const float bg_dot = bg[0] * dC[0] + bg[1] * dC[1] + bg[2] * dC[2];
        float Tc = T_final, acc[3] = {0, 0, 0}, last_alpha = 0.0f, last_color[3] = {0, 0, 0};
        int amb = 0;
        double cond = 0.0;
        for (uint32_t k = start + last[pix]; (tainted || share) && k-- > start;) {
          const float* r = staged + (size_t)(k - start) * 12;
          const float dx = r[0] - cx, dy = r[1] - cy;
          const float power = -0.5f * (r[2] * dx * dx + r[4] * dy * dy) - r[3] * dx * dy;
          const float mag = power_magnitude(r, dx, dy);
          if (clearly_faint(power, cut[k], mag)) continue;
          const float raw = r[5] * expf(power);
          amb |= near_threshold(power, raw, mag);
          const double a = raw > 0.99f ? 0.99 : (double)raw;
          if (power <= 0.0f && a >= 1.0 / 255.0) cond += a / (1.0 - a);
        }
        cond = cond * 0x1p-20 + 0x1p-18;
        for (uint32_t k = start + last[pix]; k-- > start;) {
          const float* r = staged + (size_t)(k - start) * 12;
          const float dx = r[0] - cx, dy = r[1] - cy;
          const float power = -0.5f * (r[2] * dx * dx + r[4] * dy * dy) - r[3] * dx * dy;
          if (clearly_faint(power, cut[k], power_magnitude(r, dx, dy))) continue;
          const float G = expf(power);
          const float raw = r[5] * G;
          const float alpha = fminf(0.99f, raw);
          if (amb && amb_pair) amb_pair[k] = 1; /* in or out of the sum, or weighted otherwise, elsewhere */
          const int skipped = power > 0.0f || alpha < 1.0f / 255.0f;
          /* A splat skipped here at a threshold may be composited by another
           * evaluation: its term as if composited goes into the share. */
          const int hypothetical = skipped && amb_terms && near_threshold(power, raw, power_magnitude(r, dx, dy));
          if (skipped && !hypothetical) continue;
          const float a_k = hypothetical ? fmaxf(alpha, 1.0f / 255.0f) : alpha;
          const float T_k = Tc / (1.0f - a_k);
          float acc_k[3];
          for (int c = 0; c < 3; ++c) acc_k[c] = last_alpha * last_color[c] + (1.0f - last_alpha) * acc[c];
          float dL_dalpha = 0.0f;
          for (int c = 0; c < 3; ++c) dL_dalpha += (r[6 + c] - acc_k[c]) * dC[c];
          dL_dalpha *= T_k;
          dL_dalpha += (-T_final / (1.0f - a_k)) * bg_dot;
          double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
          for (int c = 0; c < 3; ++c) v[6 + c] = (double)(a_k * T_k * dC[c]);
          if (raw < 0.99f) {
            const float dpow = raw * dL_dalpha;
            v[5] = (double)(G * dL_dalpha);
            v[0] = (double)(dpow * -(r[2] * dx + r[3] * dy));
            v[1] = (double)(dpow * -(r[4] * dy + r[3] * dx));
            v[2] = (double)(dpow * (-0.5f * dx * dx));
            v[3] = (double)(dpow * (-dx * dy));
            v[4] = (double)(dpow * (-0.5f * dy * dy));
          }
          if (amb_terms) {
            double* a = amb_terms + (size_t)k * 9;
            for (int c = 0; c < 9; ++c) a[c] += fabs(v[c]) * (amb ? 1.0 : cond);
          }
          if (hypothetical) continue;
          Tc = T_k;
          for (int c = 0; c < 3; ++c) {
            acc[c] = acc_k[c];
            last_color[c] = r[6 + c];
          }
          last_alpha = alpha;
          double* o = terms + (size_t)k * 9;
          for (int c = 0; c < 9; ++c) o[c] += v[c];
        }
      }
    }
  }
  free(staged);
  }
  memset(grad2d, 0, sizeof(double) * n * 9);
  if (tainted) memset(tainted, 0, n);
  if (share) memset(share, 0, sizeof(double) * n * 9);
  for (uint32_t tile = 0; tile < T; ++tile) {
    for (uint32_t k = ranges[2 * tile]; k < ranges[2 * tile + 1]; ++k) {
      const uint32_t i = pair_gaussians[k];
      for (int c = 0; c < 9; ++c) grad2d[(size_t)i * 9 + c] += terms[(size_t)k * 9 + c];
      if (tainted && amb_pair[k]) tainted[i] = 1;
      if (share)
        for (int c = 0; c < 9; ++c) share[(size_t)i * 9 + c] += amb_terms[(size_t)k * 9 + c];
    }
  }
  free(terms);
  free(amb_terms);
  free(amb_pair);
  free(cut);
}

/* ===================== viewer renderer (splat_preprocess.comp, splat.vert/.frag) === */
/* Round to the nearest fp16 value, ties to even, back to float: the
 * precision at which the viewer stores colour and opacity (packHalf2x16). */
float orc_half(float f) {
  uint32_t x = f2i(f);
  const uint32_t sign = x & 0x80000000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return f; /* inf / nan */
  /* fp16 step at this magnitude: 2^(e-10) for normals, 2^-24 below 2^-14 */
  int e = (int)(ax >> 23) - 127;
  if (e > 15) return i2f(sign | 0x7f800000u);
  const int step_exp = e < -14 ? -24 : e - 10;
  const double step = ldexp(1.0, step_exp);
  const double v = (double)i2f(ax) / step;
  double r = floor(v);
  const double frac = v - r;
  if (frac > 0.5 || (frac == 0.5 && fmod(r, 2.0) != 0.0)) r += 1.0;
  double out = r * step;
  if (out >= 65520.0) return i2f(sign | 0x7f800000u);
  return i2f(sign | f2i((float)out));
}

typedef struct {
  float mx, my, u1x, u1y, u2x, u2y, op, col[3];
  uint32_t lo, hi, key;
  int drawn;
} view_splat;

static void view_project(const float* p, const float* rest, int coeffs, const orc_camera* cam, uint32_t degree,
                         int antialiased, view_splat* o) {
  const float* R = cam->rotation;
  memset(o, 0, sizeof(*o));
  o->key = 0xffffu;
  const float tx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + cam->translation[0];
  const float ty = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + cam->translation[1];
  const float tz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + cam->translation[2];
  float opacity = 1.0f / (1.0f + orc_expf(-p[3]));
  if (!(p[7] > 0.5f && tz > 0.02f && opacity >= 1.0f / 255.0f)) return;
  const float fx = cam->fx, fy = cam->fy, cx = cam->cx, cy = cam->cy;
  const float W = (float)cam->width, H = (float)cam->height;
  const float inv_z = 1.0f / tz;
  const float mx = fx * tx * inv_z + cx, my = fy * ty * inv_z + cy;
  const float lim_x = 1.3f * fmaxf(cx, W - cx) / fx, lim_y = 1.3f * fmaxf(cy, H - cy) / fy;
  const float rx = fminf(fmaxf(tx * inv_z, -lim_x), lim_x), ry = fminf(fmaxf(ty * inv_z, -lim_y), lim_y);
  const float j00 = fx * inv_z, j02 = -fx * rx * inv_z, j11 = fy * inv_z, j12 = -fy * ry * inv_z;
  const float t00 = j00 * R[0] + j02 * R[6], t01 = j00 * R[1] + j02 * R[7], t02 = j00 * R[2] + j02 * R[8];
  const float t10 = j11 * R[3] + j12 * R[6], t11 = j11 * R[4] + j12 * R[7], t12 = j11 * R[5] + j12 * R[8];
  const float* q = p + 8;
  const float qlen = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), 1e-12f);
  const float w = q[0] / qlen, x = q[1] / qlen, y = q[2] / qlen, z = q[3] / qlen;
  const float sx = orc_expf(p[4]), sy = orc_expf(p[5]), sz = orc_expf(p[6]);
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
  const float ca = ra + 0.3f, cc = rc + 0.3f;
  const float det = ca * cc - rb * rb;
  if (antialiased) opacity = opacity * sqrtf(fmaxf(det_raw, 0.0f) / fmaxf(det, 1e-12f));
  if (!(det > 0.0f && opacity >= 1.0f / 255.0f)) return;
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
  if (!(mx + extent > 0.0f && mx - extent < W && my + extent > 0.0f && my - extent < H)) return;
  const float dx = p[0] - cam->position[0], dy = p[1] - cam->position[1], dz = p[2] - cam->position[2];
  const float dlen = sqrtf(dx * dx + dy * dy + dz * dz);
  float basis[16];
  sh_basis32(dx / dlen, dy / dlen, dz / dlen, degree, basis);
  float col[3] = {basis[0] * p[12], basis[0] * p[13], basis[0] * p[14]};
  int use = rest_coeffs((int)degree);
  if (use > coeffs) use = coeffs;
  for (int k = 0; k < use; ++k) {
    for (int c = 0; c < 3; ++c) col[c] = col[c] + basis[k + 1] * rest[k * 3 + c];
  }
  for (int c = 0; c < 3; ++c) col[c] = fmaxf(col[c] + 0.5f, 0.0f);
  const float n1 = a1x * a1x + a1y * a1y, n2 = a2x * a2x + a2y * a2y;
  o->mx = mx, o->my = my;
  o->u1x = a1x / n1, o->u1y = a1y / n1, o->u2x = a2x / n2, o->u2y = a2y / n2;
  o->op = orc_half(opacity);
  for (int c = 0; c < 3; ++c) o->col[c] = orc_half(col[c]);
  const float ex = 2.0f * (fabsf(a1x) + fabsf(a2x)), ey = 2.0f * (fabsf(a1y) + fabsf(a2y));
  const int x0 = (int)fmaxf(ceilf(mx - ex - 0.5f), 0.0f), x1 = (int)fminf(floorf(mx + ex - 0.5f), W - 1.0f);
  const int y0 = (int)fmaxf(ceilf(my - ey - 0.5f), 0.0f), y1 = (int)fminf(floorf(my + ey - 0.5f), H - 1.0f);
  if (x0 <= x1 && y0 <= y1) {
    o->lo = (uint32_t)x0 | ((uint32_t)y0 << 16);
    o->hi = (uint32_t)x1 | ((uint32_t)y1 << 16);
  } else {
    o->lo = 1u, o->hi = 0u;
  }
  const float u = orc_logf(tz / 0.02f) / orc_logf(1000.0f / 0.02f);
  o->key = 65534u - (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * 65534.0f);
  o->drawn = 1;
}

static int bbox_overlaps(uint32_t lo, uint32_t hi, int rx0, int ry0, int rx1, int ry1) {
  const int bx0 = (int)(lo & 0xffffu), by0 = (int)(lo >> 16), bx1 = (int)(hi & 0xffffu), by1 = (int)(hi >> 16);
  return bx0 <= rx1 && bx1 >= rx0 && by0 <= ry1 && by1 >= ry0;
}

static const view_splat* g_view_splats;
static int cmp_view(const void* a, const void* b) {
  const uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
  const uint32_t ka = g_view_splats[ia].key, kb = g_view_splats[ib].key;
  if (ka != kb) return ka < kb ? -1 : 1;
  return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

/* The viewer image: params [n][16], rest [n][coeffs][3] (used as given: pass
 * them rounded to fp16 for the reference's storage), RGBA8 out, *visible =
 * splats drawn.  Per pixel, every drawn splat in key order (ties by index)
 * whose bbox reaches the pixel's 8x8 quadrant and whose quad holds the
 * pixel centre is blended into an 8-bit target: colour = round(min(c a, 1)
 * 255 + (1 - a) colour), the same for alpha with a. */
void orc_view_render32(const float* params, const float* rest, uint32_t n, uint32_t coeffs, const orc_camera* cam,
                       uint32_t degree, int antialiased, const float clear[4], uint8_t* rgba, uint32_t* visible) {
  view_splat* v = (view_splat*)calloc(n ? n : 1, sizeof(view_splat));
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t drawn = 0;
  for (uint32_t i = 0; i < n; ++i) {
    view_project(params + (size_t)i * 16, rest + (size_t)i * coeffs * 3, (int)coeffs, cam, degree, antialiased, &v[i]);
    if (v[i].drawn) order[drawn++] = i;
  }
  g_view_splats = v;
  qsort(order, drawn, sizeof(uint32_t), cmp_view);
  for (uint32_t py = 0; py < cam->height; ++py) {
    for (uint32_t px = 0; px < cam->width; ++px) {
      const int qx = (int)(px / 8) * 8, qy = (int)(py / 8) * 8;
      const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
      float c[4];
      for (int k = 0; k < 4; ++k) c[k] = rintf(fminf(fmaxf(clear[k], 0.0f), 1.0f) * 255.0f);
      for (uint32_t t = 0; t < drawn; ++t) {
        const view_splat* s = &v[order[t]];
        if (!bbox_overlaps(s->lo, s->hi, qx, qy, qx + 7, qy + 7)) continue;
        const float dx = cx - s->mx, dy = cy - s->my;
        const float lx = dx * s->u1x + dy * s->u1y, ly = dx * s->u2x + dy * s->u2y;
        const float alpha = fminf(s->op * orc_expf(-(lx * lx + ly * ly)), 0.99f);
        if (!(fabsf(lx) <= 2.0f && fabsf(ly) <= 2.0f && alpha >= 1.0f / 255.0f)) continue;
        const float keep = 1.0f - alpha;
        for (int k = 0; k < 3; ++k) c[k] = rintf(fminf(fminf(s->col[k] * alpha, 1.0f) * 255.0f + keep * c[k], 255.0f));
        c[3] = rintf(fminf(alpha * 255.0f + keep * c[3], 255.0f));
      }
      uint8_t* o = rgba + ((size_t)py * cam->width + px) * 4;
      for (int k = 0; k < 4; ++k) o[k] = (uint8_t)c[k];
    }
  }
  *visible = drawn;
  free(order);
  free(v);
}

/* ===================== fp64 reference ====================================== */
typedef struct {
  int visible;
  double mx, my, depth, opacity;
  double con[3];
  double cov[3]; /* dilated a, b, c */
  double color[3];
  int clamped[3];
  double faint; /* ln(1 / (255 opacity)): below this power alpha < 1/255 */
} ref_t;

/* Rotation of the normalised quaternion (w, x, y, z), row-major. */
static void quat_matrix64(const double q[4], double R[3][3]) {
  const double inv = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double w = q[0] * inv, x = q[1] * inv, y = q[2] * inv, z = q[3] * inv;
  const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const double wx = w * x, wy = w * y, wz = w * z;
  const double rows[3][3] = {{1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy)},
                             {2 * (xy + wz), 1 - 2 * (xx + zz), 2 * (yz - wx)},
                             {2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy)}};
  memcpy(R, rows, sizeof(rows));
}

/* p: 16 params; sh: coeffs*3 ([coeff][rgb]).  train_selftest.cpp:79-150. */
static ref_t project64(const orc_camera* cam, const double* p, const double* sh, int degree) {
  ref_t r;
  memset(&r, 0, sizeof(r));
  double t[3];
  for (int k = 0; k < 3; ++k) {
    t[k] = cam->translation[k];
    for (int c = 0; c < 3; ++c) t[k] += (double)cam->rotation[k * 3 + c] * p[c];
  }
  const double opacity = 1.0 / (1.0 + exp(-p[3]));
  if (p[7] < 0.5 || t[2] <= NEAR || opacity < MIN_ALPHA) return r;
  const double fx = cam->fx, fy = cam->fy, cx = cam->cx, cy = cam->cy, W = cam->width, H = cam->height;
  const double z = t[2];
  const double lim_x = 1.3 * fmax(cx, W - cx) / fx, lim_y = 1.3 * fmax(cy, H - cy) / fy;
  const double tx = fmin(fmax(t[0] / z, -lim_x), lim_x), ty = fmin(fmax(t[1] / z, -lim_y), lim_y);
  const double J[2][3] = {{fx / z, 0, -fx * tx / z}, {0, fy / z, -fy * ty / z}};
  double R[3][3], M[3][3], S[3][3], T[2][3];
  quat_matrix64(p + 8, R);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) M[a][b] = R[a][b] * exp(p[4 + b]);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      S[a][b] = 0;
      for (int k = 0; k < 3; ++k) S[a][b] += M[a][k] * M[b][k];
    }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b) {
      T[a][b] = 0;
      for (int k = 0; k < 3; ++k) T[a][b] += J[a][k] * cam->rotation[k * 3 + b];
    }
  double cov[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      cov[a][b] = 0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) cov[a][b] += T[a][k] * S[k][l] * T[b][l];
    }
  const double A = cov[0][0] + DILATION, B = cov[0][1], C = cov[1][1] + DILATION;
  const double det = A * C - B * B;
  if (det <= 0) return r;
  const double inv_z = 1.0 / z, inv_det = 1.0 / det;
  r.visible = 1;
  r.depth = z;
  r.mx = cx + fx * t[0] * inv_z;
  r.my = cy + fy * t[1] * inv_z;
  r.cov[0] = A, r.cov[1] = B, r.cov[2] = C;
  r.con[0] = C * inv_det, r.con[1] = -B * inv_det, r.con[2] = A * inv_det;
  r.opacity = opacity;
  r.faint = -log(255.0 * opacity);
  double dir[3], len = 0;
  for (int k = 0; k < 3; ++k) {
    dir[k] = p[k] - cam->position[k];
    len += dir[k] * dir[k];
  }
  len = sqrt(len);
  for (int k = 0; k < 3; ++k) dir[k] /= len;
  double b[16];
  orc_sh_basis64(dir, degree, b);
  const int coeffs = rest_coeffs(degree);
  for (int c = 0; c < 3; ++c) {
    double v = b[0] * p[12 + c];
    for (int k = 0; k < coeffs; ++k) v += b[k + 1] * sh[k * 3 + c];
    v += 0.5;
    r.clamped[c] = v < 0.0;
    r.color[c] = v < 0.0 ? 0.0 : v;
  }
  return r;
}

/* out [n][16]: visible, mx, my, depth, opacity, conic abc, colour rgb, cov abc, - */
void orc_project64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                   int degree, double* out) {
  for (uint32_t i = 0; i < n; ++i) {
    const ref_t r = project64(cam, params + (size_t)i * 16, sh + (size_t)i * coeffs_in * 3, degree);
    double* o = out + (size_t)i * 16;
    o[0] = r.visible;
    o[1] = r.mx;
    o[2] = r.my;
    o[3] = r.depth;
    o[4] = r.opacity;
    o[5] = r.con[0];
    o[6] = r.con[1];
    o[7] = r.con[2];
    o[8] = r.color[0];
    o[9] = r.color[1];
    o[10] = r.color[2];
    o[11] = r.cov[0];
    o[12] = r.cov[1];
    o[13] = r.cov[2];
    o[14] = o[15] = 0.0;
  }
}

typedef struct {
  ref_t* refs;
  uint32_t* order;
  uint32_t count;
} scene64_t;

static const ref_t* g_sort_refs;
static int order_cmp(const void* a, const void* b) {
  const uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
  const double di = g_sort_refs[i].depth, dj = g_sort_refs[j].depth;
  if (di != dj) return di < dj ? -1 : 1;
  return i < j ? -1 : (i > j ? 1 : 0);
}

/* The fp64 projection of every slot, and the visible ones in depth order
 * (sorted only when `sorted`: the adjoint needs no order). */
static scene64_t project_scene64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in,
                                 const orc_camera* cam, int degree, int sorted) {
  scene64_t s;
  s.refs = (ref_t*)calloc(n > 0 ? n : 1, sizeof(ref_t));
  s.order = (uint32_t*)malloc(sizeof(uint32_t) * (n > 0 ? n : 1));
  s.count = 0;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    s.refs[i] = project64(cam, params + (size_t)i * 16, sh + (size_t)i * coeffs_in * 3, degree);
  }
  for (uint32_t i = 0; i < n; ++i)
    if (s.refs[i].visible) s.order[s.count++] = i;
  if (sorted) {
    g_sort_refs = s.refs;
    qsort(s.order, s.count, sizeof(uint32_t), order_cmp);
  }
  return s;
}

static scene64_t build_scene64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in,
                               const orc_camera* cam, int degree) {
  return project_scene64(params, sh, n, coeffs_in, cam, degree, 1);
}

static void free_scene64(scene64_t* s) {
  free(s->refs);
  free(s->order);
}

/* The pixels where a splat can reach alpha >= 1/255: its conic form q =
 * a dx^2 + 2 b dx dy + c dy^2 is at most 2 ln(255 o) there, an ellipse whose
 * half extents are sqrt(qmax A) and sqrt(qmax C) (A, C the dilated
 * covariance's diagonal); one pixel of margin.  0 when it reaches none. */
static int bbox64(const ref_t* g, uint32_t W, uint32_t H, int box[4]) {
  if (!g->visible || !(g->opacity >= MIN_ALPHA)) return 0;
  const double qmax = 2.0 * log(255.0 * g->opacity);
  if (!(qmax >= 0.0)) return 0;
  const double ex = sqrt(qmax * g->cov[0]) + 1.0, ey = sqrt(qmax * g->cov[2]) + 1.0;
  const double x0 = fmax(floor(g->mx - ex - 0.5), 0.0), x1 = fmin(ceil(g->mx + ex - 0.5), (double)W - 1.0);
  const double y0 = fmax(floor(g->my - ey - 0.5), 0.0), y1 = fmin(ceil(g->my + ey - 0.5), (double)H - 1.0);
  if (!(x0 <= x1 && y0 <= y1)) return 0;
  box[0] = (int)x0, box[1] = (int)x1, box[2] = (int)y0, box[3] = (int)y1;
  return 1;
}

/* Per 16x16 tile, the visible splats whose bbox64 meets it, in depth order:
 * list[start[t] .. start[t + 1]). */
typedef struct {
  uint32_t tiles_x, tiles_y;
  uint64_t* start;
  uint32_t* list;
} tiles64_t;

static tiles64_t bin_scene64(const scene64_t* s, uint32_t W, uint32_t H) {
  tiles64_t t;
  t.tiles_x = (W + TILE - 1) / TILE;
  t.tiles_y = (H + TILE - 1) / TILE;
  const uint32_t T = t.tiles_x * t.tiles_y;
  t.start = (uint64_t*)calloc((size_t)T + 1, sizeof(uint64_t));
  int* boxes = (int*)malloc(sizeof(int) * 4 * (s->count > 0 ? s->count : 1));
  uint8_t* hit = (uint8_t*)malloc(s->count > 0 ? s->count : 1);
  for (uint32_t k = 0; k < s->count; ++k) {
    int* b = boxes + 4 * k;
    hit[k] = (uint8_t)bbox64(&s->refs[s->order[k]], W, H, b);
    if (!hit[k]) continue;
    for (int ty = b[2] / TILE; ty <= b[3] / TILE; ++ty)
      for (int tx = b[0] / TILE; tx <= b[1] / TILE; ++tx) ++t.start[(size_t)ty * t.tiles_x + tx + 1];
  }
  for (uint32_t k = 0; k < T; ++k) t.start[k + 1] += t.start[k];
  uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)T + 1));
  memcpy(fill, t.start, sizeof(uint64_t) * ((size_t)T + 1));
  t.list = (uint32_t*)malloc(sizeof(uint32_t) * (t.start[T] > 0 ? t.start[T] : 1));
  for (uint32_t k = 0; k < s->count; ++k) {
    if (!hit[k]) continue;
    const int* b = boxes + 4 * k;
    for (int ty = b[2] / TILE; ty <= b[3] / TILE; ++ty)
      for (int tx = b[0] / TILE; tx <= b[1] / TILE; ++tx) t.list[fill[(size_t)ty * t.tiles_x + tx]++] = s->order[k];
  }
  free(boxes);
  free(hit);
  free(fill);
  return t;
}

static void free_tiles64(tiles64_t* t) {
  free(t->start);
  free(t->list);
}

/* Front-to-back compositing at one pixel over its tile's list
 * (train_selftest.cpp:243-256); returns the final T, writes C (before the
 * background) and, if asked, the list position after the last contributor. */
static double composite64(const scene64_t* s, const tiles64_t* t, uint32_t x, uint32_t y, double C[3],
                          uint64_t* end) {
  const uint32_t tile = (y / TILE) * t->tiles_x + x / TILE;
  double T = 1;
  C[0] = C[1] = C[2] = 0;
  uint64_t stop = t->start[tile];
  for (uint64_t k = t->start[tile]; k < t->start[tile + 1]; ++k) {
    const ref_t* g = &s->refs[t->list[k]];
    const double dx = g->mx - (x + 0.5), dy = g->my - (y + 0.5);
    const double power = -0.5 * (g->con[0] * dx * dx + g->con[2] * dy * dy) - g->con[1] * dx * dy;
    if (power > 0 || power < g->faint - 1e-9) continue; /* (the second: alpha < 1/255, exp not needed) */
    double alpha = g->opacity * exp(power);
    if (alpha > MAX_ALPHA) alpha = MAX_ALPHA;
    if (alpha < MIN_ALPHA) continue;
    if (T * (1 - alpha) < MIN_T) break;
    for (int c = 0; c < 3; ++c) C[c] += g->color[c] * alpha * T;
    T *= 1 - alpha;
    stop = k + 1;
  }
  if (end != NULL) *end = stop;
  return T;
}

/* Compositing of every visible gaussian at every pixel, front to back.  A
 * splat outside its bbox64 has alpha < 1/255 there and is skipped, so the
 * tile lists give the brute-force result (same operations in the same
 * order).  rgb [H*W][3] = C + T bg; T [H*W] (may be null). */
static void render_scene64(const scene64_t* s, const orc_camera* cam, const double bg[3], double* rgb, double* Tout) {
  tiles64_t t = bin_scene64(s, cam->width, cam->height);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t y = 0; y < cam->height; ++y) {
    for (uint32_t x = 0; x < cam->width; ++x) {
      double C[3];
      const double T = composite64(s, &t, x, y, C, NULL);
      const size_t pix = (size_t)y * cam->width + x;
      for (int c = 0; c < 3; ++c) rgb[pix * 3 + c] = C[c] + T * bg[c];
      if (Tout != NULL) Tout[pix] = T;
    }
  }
  free_tiles64(&t);
}

/* The brute-force loop over every visible gaussian, kept as the check of
 * the tiled render. */
void orc_render64_brute(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in,
                        const orc_camera* cam, int degree, const double bg[3], double* rgb) {
  scene64_t s = build_scene64(params, sh, n, coeffs_in, cam, degree);
  for (uint32_t y = 0; y < cam->height; ++y) {
    for (uint32_t x = 0; x < cam->width; ++x) {
      double T = 1, C[3] = {0, 0, 0};
      for (uint32_t k = 0; k < s.count; ++k) {
        const ref_t* g = &s.refs[s.order[k]];
        const double dx = g->mx - (x + 0.5), dy = g->my - (y + 0.5);
        const double power = -0.5 * (g->con[0] * dx * dx + g->con[2] * dy * dy) - g->con[1] * dx * dy;
        if (power > 0) continue;
        double alpha = g->opacity * exp(power);
        if (alpha > MAX_ALPHA) alpha = MAX_ALPHA;
        if (alpha < MIN_ALPHA) continue;
        if (T * (1 - alpha) < MIN_T) break;
        for (int c = 0; c < 3; ++c) C[c] += g->color[c] * alpha * T;
        T *= 1 - alpha;
      }
      const size_t pix = (size_t)y * cam->width + x;
      for (int c = 0; c < 3; ++c) rgb[pix * 3 + c] = C[c] + T * bg[c];
    }
  }
  free_scene64(&s);
}

void orc_render64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                  int degree, const double bg[3], double* rgb, double* T) {
  scene64_t s = build_scene64(params, sh, n, coeffs_in, cam, degree);
  render_scene64(&s, cam, bg, rgb, T);
  free_scene64(&s);
}

/* ---- SSIM (train_selftest.cpp:178-225): window sums with zero padding ---- */
static void window64(double w[11]) {
  double total = 0;
  for (int k = -5; k <= 5; ++k) total += w[k + 5] = exp(-(double)(k * k) / (2.0 * 1.5 * 1.5));
  for (int k = 0; k < 11; ++k) w[k] /= total;
}

/* Separable blur with zero padding of one plane. */
static void blur64(const double* in, double* out, double* tmp, uint32_t W, uint32_t H) {
  double w[11];
  window64(w);
#pragma omp parallel for schedule(static)
  for (uint32_t r = 0; r < H; ++r)
    for (uint32_t c = 0; c < W; ++c) {
      double s = 0;
      for (int k = -5; k <= 5; ++k) {
        const int cc = (int)c + k;
        if (cc >= 0 && cc < (int)W) s += w[k + 5] * in[(size_t)r * W + cc];
      }
      tmp[(size_t)r * W + c] = s;
    }
#pragma omp parallel for schedule(static)
  for (uint32_t r = 0; r < H; ++r)
    for (uint32_t c = 0; c < W; ++c) {
      double s = 0;
      for (int k = -5; k <= 5; ++k) {
        const int rr = (int)r + k;
        if (rr >= 0 && rr < (int)H) s += w[k + 5] * tmp[(size_t)rr * W + c];
      }
      out[(size_t)r * W + c] = s;
    }
}

/* Sum of SSIM over pixels and channels of x, y ([H*W][3]); optional
 * grad [H*W][3] = d(sum SSIM)/dx. */
double orc_ssim64(const double* x, const double* y, uint32_t W, uint32_t H, double c1, double c2, double* grad) {
  const size_t n = (size_t)W * H;
  double* in = (double*)malloc(sizeof(double) * n * 5);
  double* mu = (double*)malloc(sizeof(double) * n * 5);
  double* tmp = (double*)malloc(sizeof(double) * n);
  double* maps = (double*)malloc(sizeof(double) * n * 3);
  double* blurred = (double*)malloc(sizeof(double) * n * 3);
  double sum = 0;
  for (int c = 0; c < 3; ++c) {
    for (size_t p = 0; p < n; ++p) {
      const double a = x[p * 3 + c], b = y[p * 3 + c];
      in[p] = a;
      in[n + p] = b;
      in[2 * n + p] = a * a;
      in[3 * n + p] = b * b;
      in[4 * n + p] = a * b;
    }
    for (int q = 0; q < 5; ++q) blur64(in + q * n, mu + q * n, tmp, W, H);
    /* Each pixel's SSIM kept, then summed in pixel order: the same total for any thread count. */
    double* smap = (double*)malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n; ++p) {
      const double mx = mu[p], my = mu[n + p];
      const double vx = mu[2 * n + p] - mx * mx, vy = mu[3 * n + p] - my * my, cxy = mu[4 * n + p] - mx * my;
      const double l = 2 * mx * my + c1, cs = 2 * cxy + c2, m = mx * mx + my * my + c1, v = vx + vy + c2;
      const double s = l * cs / (m * v);
      smap[p] = s;
      if (grad != NULL) {
        /* d s / d mu_x (holding E_xx, E_xy), d s / d E_xx, d s / d E_xy */
        const double d_exy = 2 * l / (m * v);
        const double d_exx = -s / v;
        const double d_mx = 2 * my * cs / (m * v) - s * 2 * mx / m - 2 * mx * d_exx - my * d_exy;
        maps[p] = d_mx;
        maps[n + p] = d_exx;
        maps[2 * n + p] = d_exy;
      }
    }
    for (size_t p = 0; p < n; ++p) sum += smap[p];
    free(smap);
    if (grad != NULL) {
      /* The window is symmetric: the adjoint of the blur is the blur. */
      for (int q = 0; q < 3; ++q) blur64(maps + q * n, blurred + q * n, tmp, W, H);
      for (size_t p = 0; p < n; ++p) {
        grad[p * 3 + c] = blurred[p] + 2 * x[p * 3 + c] * blurred[n + p] + y[p * 3 + c] * blurred[2 * n + p];
      }
    }
  }
  free(in);
  free(mu);
  free(tmp);
  free(maps);
  free(blurred);
  return sum;
}

/* Loss of an image x against y, both [H*W][3]:
 * (1 - lambda) sum|e| / 3HW + lambda (1 - sum SSIM / 3HW); optional dL/dx. */
double orc_loss_image64(const double* x, const double* y, uint32_t W, uint32_t H, double lambda, double* grad) {
  const size_t n = (size_t)W * H;
  const double scale = 1.0 / (3.0 * (double)n);
  double l1 = 0;
  for (size_t i = 0; i < n * 3; ++i) l1 += fabs(x[i] - y[i]);
  double loss = (1 - lambda) * l1 * scale;
  double* sg = NULL;
  if (lambda != 0.0) {
    if (grad != NULL) sg = (double*)malloc(sizeof(double) * n * 3);
    const double ssim = orc_ssim64(x, y, W, H, 0.01 * 0.01, 0.03 * 0.03, sg);
    loss += lambda * (1.0 - ssim * scale);
  }
  if (grad != NULL) {
    for (size_t i = 0; i < n * 3; ++i) {
      const double e = x[i] - y[i];
      grad[i] = (e > 0 ? 1.0 : (e < 0 ? -1.0 : 0.0)) * (1 - lambda) * scale;
      if (sg != NULL) grad[i] -= lambda * scale * sg[i];
    }
  }
  free(sg);
  return loss;
}

/* Target [H*W] RGBA8 -> [H*W][3] double (masked: over bg by alpha). */
static void target64(const uint8_t* rgba, uint32_t W, uint32_t H, const double bg[3], int masked, double* out) {
  for (size_t p = 0; p < (size_t)W * H; ++p) {
    const double a = rgba[p * 4 + 3] / 255.0;
    for (int c = 0; c < 3; ++c) {
      const double t = rgba[p * 4 + c] / 255.0;
      out[p * 3 + c] = masked ? bg[c] * (1 - a) + t * a : t;
    }
  }
}

double orc_loss64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                  int degree, const double bg[3], const uint8_t* target, int masked, double lambda) {
  const size_t pixels = (size_t)cam->width * cam->height;
  double* rgb = (double*)malloc(sizeof(double) * pixels * 3);
  double* tgt = (double*)malloc(sizeof(double) * pixels * 3);
  orc_render64(params, sh, n, coeffs_in, cam, degree, bg, rgb, NULL);
  target64(target, cam->width, cam->height, bg, masked, tgt);
  const double loss = orc_loss_image64(rgb, tgt, cam->width, cam->height, lambda, NULL);
  free(rgb);
  free(tgt);
  return loss;
}

/* Projection adjoint (train_step.glsl:102-219), per gaussian: dL/d(2D splat)
 * grad2d [n][9] -> dL/d(16 params) gparams and dL/d(SH rest) gsh. */
static void adjoint_scene64(const scene64_t* sc, const double* params, const double* sh, uint32_t n,
                            uint32_t coeffs_in, const orc_camera* cam, int degree, const double* grad2d,
                            double* gparams, double* gsh, double* pose) {
  memset(gparams, 0, sizeof(double) * n * 16);
  if (pose != NULL) memset(pose, 0, sizeof(double) * n * 6);
  if (coeffs_in > 0) memset(gsh, 0, sizeof(double) * n * coeffs_in * 3);
  const double* Rc = NULL;
  double V[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) V[r][c] = cam->rotation[r * 3 + c];
  (void)Rc;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    const ref_t* g = &sc->refs[i];
    if (!g->visible) continue;
    const double* p = params + (size_t)i * 16;
    const double* sp = sh + (size_t)i * coeffs_in * 3;
    const double* o = grad2d + (size_t)i * 9;
    double* gp = gparams + (size_t)i * 16;
    double t[3];
    for (int k = 0; k < 3; ++k) t[k] = cam->translation[k] + V[k][0] * p[0] + V[k][1] * p[1] + V[k][2] * p[2];
    const double fx = cam->fx, fy = cam->fy, cx = cam->cx, cy = cam->cy;
    const double Wd = cam->width, Hd = cam->height, z = t[2];
    const double lim_x = 1.3 * fmax(cx, Wd - cx) / fx, lim_y = 1.3 * fmax(cy, Hd - cy) / fy;
    const double rx = t[0] / z, ry = t[1] / z;
    const int clamp_x = rx < -lim_x || rx > lim_x, clamp_y = ry < -lim_y || ry > lim_y;
    const double ux = fmin(fmax(rx, -lim_x), lim_x), uy = fmin(fmax(ry, -lim_y), lim_y);
    const double J[2][3] = {{fx / z, 0, -fx * ux / z}, {0, fy / z, -fy * uy / z}};
    double T[2][3];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) T[a][b] = J[a][0] * V[0][b] + J[a][1] * V[1][b] + J[a][2] * V[2][b];
    const double* q = p + 8;
    const double ql = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double qn[4] = {q[0] / ql, q[1] / ql, q[2] / ql, q[3] / ql};
    double R[3][3], M[3][3], S[3][3];
    quat_matrix64(q, R);
    const double sc[3] = {exp(p[4]), exp(p[5]), exp(p[6])};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) M[a][b] = R[a][b] * sc[b];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) S[a][b] = M[a][0] * M[b][0] + M[a][1] * M[b][1] + M[a][2] * M[b][2];
    const double A = g->cov[0], B = g->cov[1], C = g->cov[2], D = A * C - B * B;
    /* colour */
    double dir[3], len = 0;
    for (int k = 0; k < 3; ++k) {
      dir[k] = p[k] - cam->position[k];
      len += dir[k] * dir[k];
    }
    len = sqrt(len);
    for (int k = 0; k < 3; ++k) dir[k] /= len;
    double dcol[3];
    for (int c = 0; c < 3; ++c) dcol[c] = g->clamped[c] ? 0.0 : o[6 + c];
    double basis[16];
    orc_sh_basis64(dir, degree, basis);
    for (int c = 0; c < 3; ++c) gp[12 + c] = basis[0] * dcol[c];
    const int coeffs = rest_coeffs(degree) < (int)coeffs_in ? rest_coeffs(degree) : (int)coeffs_in;
    for (int k = 0; k < coeffs; ++k)
      for (int c = 0; c < 3; ++c) gsh[((size_t)i * coeffs_in + k) * 3 + c] = basis[k + 1] * dcol[c];
    /* direction gradient by central differences of the basis is avoided: the
       analytic basis derivative, written out from the polynomials above. */
    double ddir[3] = {0, 0, 0};
    {
      const double x = dir[0], y = dir[1], zz_ = dir[2];
      double gb[16][3];
      memset(gb, 0, sizeof(gb));
      const double C1 = 0.4886025119029199;
      gb[1][1] = -C1;
      gb[2][2] = C1;
      gb[3][0] = -C1;
      const double xx = x * x, yy = y * y, zz = zz_ * zz_;
      const double Aa = 1.0925484305920792, Bb = 0.31539156525252005, Cc = 0.5462742152960396;
      gb[4][0] = Aa * y, gb[4][1] = Aa * x;
      gb[5][1] = -Aa * zz_, gb[5][2] = -Aa * y;
      gb[6][0] = -2 * Bb * x, gb[6][1] = -2 * Bb * y, gb[6][2] = 4 * Bb * zz_;
      gb[7][0] = -Aa * zz_, gb[7][2] = -Aa * x;
      gb[8][0] = 2 * Cc * x, gb[8][1] = -2 * Cc * y;
      const double Dd = -0.5900435899266435, Ee = 2.890611442640554, Ff = -0.4570457994644658;
      const double Gg = 0.3731763325901154, Hh = 1.445305721320277;
      gb[9][0] = Dd * 6 * x * y, gb[9][1] = Dd * (3 * xx - 3 * yy);
      gb[10][0] = Ee * y * zz_, gb[10][1] = Ee * x * zz_, gb[10][2] = Ee * x * y;
      gb[11][0] = Ff * (-2 * x * y), gb[11][1] = Ff * (4 * zz - xx - 3 * yy), gb[11][2] = Ff * 8 * y * zz_;
      gb[12][0] = Gg * (-6 * x * zz_), gb[12][1] = Gg * (-6 * y * zz_), gb[12][2] = Gg * (6 * zz - 3 * xx - 3 * yy);
      gb[13][0] = Ff * (4 * zz - 3 * xx - yy), gb[13][1] = Ff * (-2 * x * y), gb[13][2] = Ff * 8 * x * zz_;
      gb[14][0] = Hh * 2 * x * zz_, gb[14][1] = Hh * (-2 * y * zz_), gb[14][2] = Hh * (xx - yy);
      gb[15][0] = Dd * (3 * xx - 3 * yy), gb[15][1] = Dd * (-6 * x * y);
      for (int k = 0; k < coeffs; ++k) {
        double w = 0;
        for (int c = 0; c < 3; ++c) w += sp[k * 3 + c] * dcol[c];
        for (int d = 0; d < 3; ++d) ddir[d] += w * gb[k + 1][d];
      }
    }
    const double dd = dir[0] * ddir[0] + dir[1] * ddir[1] + dir[2] * ddir[2];
    double dpos[3];
    for (int k = 0; k < 3; ++k) dpos[k] = (ddir[k] - dir[k] * dd) / len;
    /* opacity */
    gp[3] = o[5] * g->opacity * (1 - g->opacity);
    /* conic -> covariance */
    const double ga = o[2], gbc = o[3], gc = o[4];
    const double gA = (-C * C * ga + B * C * gbc - B * B * gc) / (D * D);
    const double gB = (2 * B * C * ga - (D + 2 * B * B) * gbc + 2 * A * B * gc) / (D * D);
    const double gC = (-B * B * ga + A * B * gbc - A * A * gc) / (D * D);
    const double G2[2][2] = {{gA, 0.5 * gB}, {0.5 * gB, gC}};
    double dS[3][3], dT[2][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        dS[a][b] = 0;
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) dS[a][b] += T[r][a] * G2[r][c] * T[c][b];
      }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) {
        dT[a][b] = 0;
        for (int c = 0; c < 2; ++c)
          for (int k = 0; k < 3; ++k) dT[a][b] += 2 * G2[a][c] * T[c][k] * S[k][b];
      }
    double dJ[2][3];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) dJ[a][b] = dT[a][0] * V[b][0] + dT[a][1] * V[b][1] + dT[a][2] * V[b][2];
    double dt[3] = {0, 0, 0};
    const double iz2 = 1 / (z * z);
    dt[2] += -fx * iz2 * dJ[0][0] - fy * iz2 * dJ[1][1];
    if (clamp_x) {
      dt[2] += fx * ux * iz2 * dJ[0][2];
    } else {
      dt[0] += -fx * iz2 * dJ[0][2];
      dt[2] += 2 * fx * t[0] * iz2 / z * dJ[0][2];
    }
    if (clamp_y) {
      dt[2] += fy * uy * iz2 * dJ[1][2];
    } else {
      dt[1] += -fy * iz2 * dJ[1][2];
      dt[2] += 2 * fy * t[1] * iz2 / z * dJ[1][2];
    }
    dt[0] += o[0] * fx / z;
    dt[1] += o[1] * fy / z;
    dt[2] += -(o[0] * fx * t[0] + o[1] * fy * t[1]) * iz2;
    for (int k = 0; k < 3; ++k) gp[k] = V[0][k] * dt[0] + V[1][k] * dt[1] + V[2][k] * dt[2] + dpos[k];
    if (pose != NULL) {
      /* Camera pose (train_step.glsl:196-201): a shift d moves t by d; a turn
         w moves t by w x t and turns W in T = J W by [w]x, whose share is the
         antisymmetric part of GW = J^T dT W^T.  The view direction's share
         (colour through SH) is left out, as in the reference. */
      double P[3][3], GW[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P[r][c] = J[0][r] * dT[0][c] + J[1][r] * dT[1][c];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) GW[r][c] = P[r][0] * V[c][0] + P[r][1] * V[c][1] + P[r][2] * V[c][2];
      double* ps = pose + (size_t)i * 6;
      for (int k = 0; k < 3; ++k) ps[k] = dt[k];
      ps[3] = t[1] * dt[2] - t[2] * dt[1] + GW[2][1] - GW[1][2];
      ps[4] = t[2] * dt[0] - t[0] * dt[2] + GW[0][2] - GW[2][0];
      ps[5] = t[0] * dt[1] - t[1] * dt[0] + GW[1][0] - GW[0][1];
    }
    /* scale and rotation */
    double dM[3][3], gR[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) dM[a][b] = 2 * (dS[a][0] * M[0][b] + dS[a][1] * M[1][b] + dS[a][2] * M[2][b]);
    for (int b = 0; b < 3; ++b) gp[4 + b] = (dM[0][b] * R[0][b] + dM[1][b] * R[1][b] + dM[2][b] * R[2][b]) * sc[b];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) gR[a][b] = dM[a][b] * sc[b];
    const double w = qn[0], x = qn[1], y = qn[2], zq = qn[3];
    const double dqn[4] = {
        2 * (-zq * gR[0][1] + y * gR[0][2] + zq * gR[1][0] - x * gR[1][2] - y * gR[2][0] + x * gR[2][1]),
        2 * (y * gR[0][1] + zq * gR[0][2] + y * gR[1][0] - 2 * x * gR[1][1] - w * gR[1][2] + zq * gR[2][0] +
             w * gR[2][1] - 2 * x * gR[2][2]),
        2 * (-2 * y * gR[0][0] + x * gR[0][1] + w * gR[0][2] + x * gR[1][0] + zq * gR[1][2] - w * gR[2][0] +
             zq * gR[2][1] - 2 * y * gR[2][2]),
        2 * (-2 * zq * gR[0][0] - w * gR[0][1] + x * gR[0][2] + w * gR[1][0] - 2 * zq * gR[1][1] + y * gR[1][2] +
             x * gR[2][0] + y * gR[2][1])};
    const double qd = qn[0] * dqn[0] + qn[1] * dqn[1] + qn[2] * dqn[2] + qn[3] * dqn[3];
    for (int k = 0; k < 4; ++k) gp[8 + k] = (dqn[k] - qn[k] * qd) / ql;
  }
}

void orc_adjoint64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                   int degree, const double* grad2d, double* gparams, double* gsh) {
  scene64_t s = project_scene64(params, sh, n, coeffs_in, cam, degree, 0);
  adjoint_scene64(&s, params, sh, n, coeffs_in, cam, degree, grad2d, gparams, gsh, NULL);
  free_scene64(&s);
}

/* Camera pose gradient per gaussian (train_step.glsl:196-201): pose [n][6] =
 * its share of dL/dd xyz, dL/dw xyz (camera axes) from grad2d [n][9]; the
 * view's gradient is the sum over the gaussians (trainer.cpp:1319-1330). */
void orc_pose_adjoint64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                        int degree, const double* grad2d, double* pose) {
  scene64_t s = build_scene64(params, sh, n, coeffs_in, cam, degree);
  double* gp = (double*)malloc(sizeof(double) * (n > 0 ? n : 1) * 16);
  double* gs = (double*)malloc(sizeof(double) * (n > 0 ? n : 1) * (coeffs_in > 0 ? coeffs_in : 1) * 3);
  adjoint_scene64(&s, params, sh, n, coeffs_in, cam, degree, grad2d, gp, gs, pose);
  free(gp);
  free(gs);
  free_scene64(&s);
}

/* ---- analytic backward ---------------------------------------------------- */
/* Per gaussian dL/d(mean xy, conic abc, opacity, colour rgb) of the loss of a
 * brute-force render against target (grad2d [n][9]), and the gradients of the
 * 16 parameters (gparams [n][16]) and SH rest (gsh [n][coeffs_in][3]).
 * Returns the loss. */
double orc_backward64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                      int degree, const double bg[3], const uint8_t* target, int masked, double lambda,
                      double* grad2d, double* gparams, double* gsh) {
  const uint32_t W = cam->width, H = cam->height;
  const size_t pixels = (size_t)W * H;
  scene64_t s = build_scene64(params, sh, n, coeffs_in, cam, degree);
  double* rgb = (double*)malloc(sizeof(double) * pixels * 3);
  double* Tf = (double*)malloc(sizeof(double) * pixels);
  double* tgt = (double*)malloc(sizeof(double) * pixels * 3);
  double* dldc = (double*)malloc(sizeof(double) * pixels * 3);
  render_scene64(&s, cam, bg, rgb, Tf);
  target64(target, W, H, bg, masked, tgt);
  const double loss = orc_loss_image64(rgb, tgt, W, H, lambda, dldc);
  memset(grad2d, 0, sizeof(double) * n * 9);
  uint32_t* contrib = (uint32_t*)malloc(sizeof(uint32_t) * (s.count > 0 ? s.count : 1));
  tiles64_t tl = bin_scene64(&s, W, H);
  for (uint32_t y = 0; y < H; ++y) {
    for (uint32_t x = 0; x < W; ++x) {
      const size_t pix = (size_t)y * W + x;
      /* Forward again: the contributors of this pixel, front to back. */
      uint32_t m = 0;
      double T = 1;
      const uint32_t tile = (y / TILE) * tl.tiles_x + x / TILE;
      for (uint64_t k = tl.start[tile]; k < tl.start[tile + 1]; ++k) {
        const ref_t* g = &s.refs[tl.list[k]];
        const double dx = g->mx - (x + 0.5), dy = g->my - (y + 0.5);
        const double power = -0.5 * (g->con[0] * dx * dx + g->con[2] * dy * dy) - g->con[1] * dx * dy;
        if (power > 0) continue;
        double alpha = g->opacity * exp(power);
        if (alpha > MAX_ALPHA) alpha = MAX_ALPHA;
        if (alpha < MIN_ALPHA) continue;
        if (T * (1 - alpha) < MIN_T) break;
        T *= 1 - alpha;
        contrib[m++] = tl.list[k];
      }
      const double* dC = dldc + pix * 3;
      const double T_final = Tf[pix];
      const double bg_dot = bg[0] * dC[0] + bg[1] * dC[1] + bg[2] * dC[2];
      /* Back to front (train_common.glsl:228-253). */
      double Tc = T_final, acc[3] = {0, 0, 0}, last_alpha = 0, last_color[3] = {0, 0, 0};
      for (uint32_t k = m; k-- > 0;) {
        const uint32_t i = contrib[k];
        const ref_t* g = &s.refs[i];
        const double dx = g->mx - (x + 0.5), dy = g->my - (y + 0.5);
        const double power = -0.5 * (g->con[0] * dx * dx + g->con[2] * dy * dy) - g->con[1] * dx * dy;
        const double G = exp(power);
        const double raw = g->opacity * G;
        const double alpha = raw > MAX_ALPHA ? MAX_ALPHA : raw;
        Tc = Tc / (1 - alpha);
        for (int c = 0; c < 3; ++c) acc[c] = last_alpha * last_color[c] + (1 - last_alpha) * acc[c];
        double dL_dalpha = 0;
        for (int c = 0; c < 3; ++c) {
          last_color[c] = g->color[c];
          dL_dalpha += (g->color[c] - acc[c]) * dC[c];
        }
        dL_dalpha *= Tc;
        last_alpha = alpha;
        dL_dalpha += (-T_final / (1 - alpha)) * bg_dot;
        double* o = grad2d + (size_t)i * 9;
        for (int c = 0; c < 3; ++c) o[6 + c] += alpha * Tc * dC[c];
        if (raw < MAX_ALPHA) {
          const double dpow = raw * dL_dalpha;
          o[5] += G * dL_dalpha;
          o[0] += dpow * -(g->con[0] * dx + g->con[1] * dy);
          o[1] += dpow * -(g->con[2] * dy + g->con[1] * dx);
          o[2] += dpow * (-0.5 * dx * dx);
          o[3] += dpow * (-dx * dy);
          o[4] += dpow * (-0.5 * dy * dy);
        }
      }
    }
  }
  free(contrib);
  free_tiles64(&tl);
  adjoint_scene64(&s, params, sh, n, coeffs_in, cam, degree, grad2d, gparams, gsh, NULL);
  free(rgb);
  free(Tf);
  free(tgt);
  free(dldc);
  free_scene64(&s);
  return loss;
}

/* ---- central differences with the compositor's branches frozen ---------- */
/* Sum over pixels of [x0, x1] x [y0, y1] (and channels) of SSIM of x
 * against y, both full [H*W][3] images, with the zero padding of the full
 * image (window sums as orc_ssim64). */
static double ssim_region64(const double* x, const double* y, uint32_t W, uint32_t H, int x0, int x1, int y0, int y1,
                            double c1, double c2) {
  double w[11];
  window64(w);
  double sum = 0;
  for (int py = y0; py <= y1; ++py) {
    for (int px = x0; px <= x1; ++px) {
      for (int c = 0; c < 3; ++c) {
        double m[5] = {0, 0, 0, 0, 0};
        for (int dy = -5; dy <= 5; ++dy) {
          const int yy = py + dy;
          if (yy < 0 || yy >= (int)H) continue;
          double r[5] = {0, 0, 0, 0, 0};
          for (int dx = -5; dx <= 5; ++dx) {
            const int xx = px + dx;
            if (xx < 0 || xx >= (int)W) continue;
            const double a = x[((size_t)yy * W + xx) * 3 + c], b = y[((size_t)yy * W + xx) * 3 + c];
            const double wx = w[dx + 5];
            r[0] += wx * a, r[1] += wx * b, r[2] += wx * a * a, r[3] += wx * b * b, r[4] += wx * a * b;
          }
          for (int q = 0; q < 5; ++q) m[q] += w[dy + 5] * r[q];
        }
        const double vx = m[2] - m[0] * m[0], vy = m[3] - m[1] * m[1], cxy = m[4] - m[0] * m[1];
        sum += (2 * m[0] * m[1] + c1) * (2 * cxy + c2) / ((m[0] * m[0] + m[1] * m[1] + c1) * (vx + vy + c2));
      }
    }
  }
  return sum;
}

/* Central differences (step h) of the fp64 loss (orc_loss64 with an
 * unmasked target: (1 - lambda) mean |e| + lambda (1 - mean SSIM)) by single
 * parameters, with every branch of the compositor frozen as it went in the
 * unperturbed render: which splats a pixel composites (alpha >= 1/255 and
 * power <= 0), where alpha is clamped at 0.99, where the walk stops
 * (T < 1e-4), which colour channels are clamped at 0, and the depth order.
 * That is the function whose derivative the analytic backward computes: it
 * leaves out the jumps where a splat's 1/255 contour crosses a pixel centre,
 * which at 1280x720 a step of 1e-4 in a position makes likely for most
 * splats and which would swamp the difference quotient.  Only the pixels the
 * perturbed gaussian composites change, so the L1 part is summed over its
 * bbox64 and the SSIM part over that box grown by the 5-pixel window radius.
 * Samples: gaussian[k], column[k] (0..15: the parameter slot; 16 + 3 j + c:
 * SH rest coefficient j, channel c).  fd[k] is NaN where the perturbation
 * makes the gaussian invisible.  Returns the unperturbed loss. */
double orc_fd_frozen64(const double* params, const double* sh, uint32_t n, uint32_t coeffs_in, const orc_camera* cam,
                       int degree, const double bg[3], const uint8_t* target, double lambda, uint32_t count,
                       const uint32_t* gaussian, const uint32_t* column, double h, double* fd) {
  const uint32_t W = cam->width, H = cam->height;
  const size_t pixels = (size_t)W * H;
  const double c1 = 0.01 * 0.01, c2 = 0.03 * 0.03, norm = 1.0 / (3.0 * (double)pixels);
  scene64_t s = build_scene64(params, sh, n, coeffs_in, cam, degree);
  tiles64_t tl = bin_scene64(&s, W, H);
  double* x = (double*)malloc(sizeof(double) * pixels * 3);
  double* y = (double*)malloc(sizeof(double) * pixels * 3);
  uint64_t* stop = (uint64_t*)malloc(sizeof(uint64_t) * pixels);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t py = 0; py < H; ++py)
    for (uint32_t px = 0; px < W; ++px) {
      double C[3];
      const size_t pix = (size_t)py * W + px;
      const double T = composite64(&s, &tl, px, py, C, &stop[pix]);
      for (int c = 0; c < 3; ++c) x[pix * 3 + c] = C[c] + T * bg[c];
    }
  target64(target, W, H, bg, 0, y);
  double* ximg = (double*)malloc(sizeof(double) * pixels * 3);
  memcpy(ximg, x, sizeof(double) * pixels * 3);
  const double base = orc_loss_image64(ximg, y, W, H, lambda, NULL);
  free(ximg);
#pragma omp parallel
  {
    double* xp = (double*)malloc(sizeof(double) * pixels * 3);
    memcpy(xp, x, sizeof(double) * pixels * 3);
    double* row = (double*)malloc(sizeof(double) * 16 + sizeof(double) * (coeffs_in > 0 ? coeffs_in : 1) * 3);
    double* rest = row + 16;
#pragma omp for schedule(dynamic, 1)
    for (uint32_t k = 0; k < count; ++k) {
      const uint32_t i = gaussian[k], col = column[k];
      const ref_t* g0 = &s.refs[i];
      int box[4];
      fd[k] = NAN;
      if (i >= n || !bbox64(g0, W, H, box)) {
        fd[k] = 0.0; /* composites nowhere: the frozen loss does not depend on it */
        continue;
      }
      const int sx0 = box[0] - 5 < 0 ? 0 : box[0] - 5, sx1 = box[1] + 5 >= (int)W ? (int)W - 1 : box[1] + 5;
      const int sy0 = box[2] - 5 < 0 ? 0 : box[2] - 5, sy1 = box[3] + 5 >= (int)H ? (int)H - 1 : box[3] + 5;
      double delta[2];
      int ok = 1;
      for (int side = 0; side < 2 && ok; ++side) {
        memcpy(row, params + (size_t)i * 16, sizeof(double) * 16);
        if (coeffs_in > 0) memcpy(rest, sh + (size_t)i * coeffs_in * 3, sizeof(double) * coeffs_in * 3);
        const double step = side == 0 ? h : -h;
        if (col < 16) row[col] += step;
        else rest[col - 16] += step;
        const ref_t g1 = project64(cam, row, rest, degree);
        if (!g1.visible) {
          ok = 0;
          break;
        }
        double l1 = 0, ssim_new, ssim_old;
        for (int py = box[2]; py <= box[3]; ++py) {
          for (int px = box[0]; px <= box[1]; ++px) {
            const uint32_t tile = ((uint32_t)py / TILE) * tl.tiles_x + (uint32_t)px / TILE;
            double T = 1, C[3] = {0, 0, 0};
            const size_t pix = (size_t)py * W + px;
            for (uint64_t q = tl.start[tile]; q < stop[pix]; ++q) {
              const uint32_t j = tl.list[q];
              const ref_t* g = &s.refs[j];
              const double dx = g->mx - (px + 0.5), dy = g->my - (py + 0.5);
              const double power = -0.5 * (g->con[0] * dx * dx + g->con[2] * dy * dy) - g->con[1] * dx * dy;
              if (power > 0) continue;
              const double raw = g->opacity * exp(power);
              if (raw < MIN_ALPHA) continue; /* (a clamped raw is >= 0.99: never below 1/255) */
              double alpha = raw > MAX_ALPHA ? MAX_ALPHA : raw;
              const double* col3 = g->color;
              if (j == i) { /* the perturbed splat, on the branches it took unperturbed */
                if (raw < MAX_ALPHA) {
                  const double ex = g1.mx - (px + 0.5), ey = g1.my - (py + 0.5);
                  const double pw = -0.5 * (g1.con[0] * ex * ex + g1.con[2] * ey * ey) - g1.con[1] * ex * ey;
                  alpha = g1.opacity * exp(pw);
                }
                col3 = g1.color; /* clamped channels: see below */
              }
              for (int c = 0; c < 3; ++c) {
                const double cv = (j == i && g->clamped[c]) ? 0.0 : col3[c];
                C[c] += cv * alpha * T;
              }
              T *= 1 - alpha;
            }
            for (int c = 0; c < 3; ++c) {
              const double v = C[c] + T * bg[c];
              l1 += fabs(v - y[pix * 3 + c]) - fabs(x[pix * 3 + c] - y[pix * 3 + c]);
              xp[pix * 3 + c] = v;
            }
          }
        }
        ssim_new = lambda != 0.0 ? ssim_region64(xp, y, W, H, sx0, sx1, sy0, sy1, c1, c2) : 0.0;
        for (int py = box[2]; py <= box[3]; ++py)
          memcpy(xp + ((size_t)py * W + box[0]) * 3, x + ((size_t)py * W + box[0]) * 3,
                 sizeof(double) * 3 * (size_t)(box[1] - box[0] + 1));
        ssim_old = lambda != 0.0 ? ssim_region64(xp, y, W, H, sx0, sx1, sy0, sy1, c1, c2) : 0.0;
        delta[side] = (1 - lambda) * l1 * norm - lambda * (ssim_new - ssim_old) * norm;
      }
      if (ok) fd[k] = (delta[0] - delta[1]) / (2 * h);
    }
    free(xp);
    free(row);
  }
  free(x);
  free(y);
  free(stop);
  free_tiles64(&tl);
  free_scene64(&s);
  return base;
}

/* ---- Adam (train_step.glsl:62-68) ---------------------------------------- */
void orc_adam(float* p, const float* g, float* m, float* v, const float* lr, uint64_t count, float beta1, float beta2,
              float eps, float bias1, float bias2) {
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; ++i) {
    m[i] = g[i] * (1.0f - beta1) + m[i] * beta1;
    v[i] = g[i] * g[i] * (1.0f - beta2) + v[i] * beta2;
    p[i] = p[i] - lr[i] * (m[i] * bias1) / (sqrtf(v[i] * bias2) + eps);
  }
}

/* ---- densify decision (train_densify_mark.comp:28-53) -------------------- */
/* action: 0 none, 1 clone, 2 split; release: 1 free or pruned; prune: 1 pruned. */
void orc_densify_mark(const float* params, const float* stats, uint32_t n, float grad_threshold,
                      float clone_max_scale, float prune_opacity, float prune_max_scale, float prune_max_radius,
                      uint32_t* action, uint32_t* release, uint32_t* prune) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = params + (size_t)i * 16;
    const float* st = stats + (size_t)i * 4;
    action[i] = 0;
    prune[i] = 0;
    release[i] = p[7] < 0.5f;
    if (release[i]) continue;
    const float max_scale = orc_expf(fmaxf(p[4], fmaxf(p[5], p[6])));
    int pr = 1.0f / (1.0f + orc_expf(-p[3])) < prune_opacity;
    if (prune_max_scale > 0 && max_scale > prune_max_scale) pr = 1;
    if (prune_max_radius > 0 && st[2] > prune_max_radius) pr = 1;
    if (pr) {
      prune[i] = 1;
      release[i] = 1;
    } else if (st[1] > 0 && st[0] / st[1] >= grad_threshold) {
      action[i] = max_scale <= clone_max_scale ? 1 : 2;
    }
  }
}
