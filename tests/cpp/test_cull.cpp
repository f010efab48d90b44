// The rasterizers' quadrant cull (csrc/cull.h, the same source the kernels
// compile) against per-pixel evaluation, without a GPU.  For needle splats
// (major-axis variance up to 1e6 px^2, minor 0.3-4 px^2) and 8x8 boxes of
// pixel centres along the major axis and across the footprint: whenever some
// pixel of a box draws the splat (o exp(power) >= 1/255, power evaluated in
// fp32 in the rasterizers' order), the cull must keep the box.  The same
// boxes under a slack without the rounding term lose drawn pixels, which
// shows these cases reach the rounding regime.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "cull.h"
#include "test_util.h"

namespace {

struct Splat {
  float mx, my, a, b, c, o;
};

// fp32 power at a pixel centre in the rasterizers' order (raster.hip eval_splat).
bool pixel_draws(const Splat& s, float cx, float cy) {
  const float A = -0.5f * s.a, B = -s.b, C = -0.5f * s.c;
  const float dx = s.mx - cx, dy = s.my - cy;
  const float adx = A * dx, cdy = C * dy;
  const float u = std::fmaf(B, dy, adx);
  const float power = std::fmaf(dx, u, cdy * dy);
  if (power > 0.0f) return false;
  return s.o * std::exp(power) >= 1.0f / 255.0f;
}

}  // namespace

int main() {
  std::mt19937_64 rng(4242);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long boxes = 0, drawn_boxes = 0, culled = 0, missed = 0, missed_old = 0, missed_rcp = 0;
  for (int n = 0; n < 4000; ++n) {
    const double major = std::exp(std::log(1e2) + U(rng) * (std::log(1e6) - std::log(1e2)));
    const double minor = 0.3 + 3.7 * U(rng);
    const double th = U(rng) * 3.141592653589793;
    const double cs = std::cos(th), sn = std::sin(th);
    const double c00 = major * cs * cs + minor * sn * sn, c11 = major * sn * sn + minor * cs * cs;
    const double c01 = (major - minor) * cs * sn;
    const double det = c00 * c11 - c01 * c01;
    Splat s;
    s.a = float(c11 / det), s.b = float(-c01 / det), s.c = float(c00 / det);
    s.mx = float(2000.0 + 16.0 * U(rng)), s.my = float(2000.0 + 16.0 * U(rng));
    s.o = float(0.01 + 0.98 * U(rng));
    const float limit = std::log(255.0f * s.o);
    const double reach = 3.2 * std::sqrt(major);
    for (int k = 0; k < 48; ++k) {
      // Along the major axis out to past the 1/255 contour, or anywhere in the bbox.
      double px, py;
      if (k < 40) {
        const double t = (k < 20 ? 1.0 : -1.0) * reach * (0.5 + 0.5 * U(rng));
        px = s.mx + t * cs - 8.0 * U(rng), py = s.my + t * sn - 8.0 * U(rng);
      } else {
        px = s.mx + reach * (2.0 * U(rng) - 1.0), py = s.my + reach * (2.0 * U(rng) - 1.0);
      }
      const int qx = int(std::floor(px)), qy = int(std::floor(py));
      bool draws = false;
      for (int j = 0; j < 64 && !draws; ++j) draws = pixel_draws(s, float(qx + (j & 7)) + 0.5f, float(qy + (j >> 3)) + 0.5f);
      const float x0 = float(qx) + 0.5f, y0 = float(qy) + 0.5f;
      const float u0 = s.mx - (x0 + 7.0f), u1 = s.mx - x0, v0 = s.my - (y0 + 7.0f), v1 = s.my - y0;
      const bool keep = gsh::quadrant_may_draw(s.a, s.b, s.c, 1.0f / s.a, 1.0f / s.c, limit, u0, u1, v0, v1);
      // The kernels' reciprocal is within an ulp or two of 1/a.
      const bool keep_rcp = gsh::quadrant_may_draw(s.a, s.b, s.c, (1.0f / s.a) * (1.0f + 0x1p-22f),
                                                   (1.0f / s.c) * (1.0f - 0x1p-22f), limit, u0, u1, v0, v1);
      const float m = gsh::box_min_quadratic(s.a, s.b, s.c, 1.0f / s.a, 1.0f / s.c, u0, u1, v0, v1);
      const bool keep_old = m <= limit + 0.02f + 0.01f * std::fabs(limit);
      ++boxes;
      drawn_boxes += draws;
      culled += !keep;
      missed += draws && !keep;
      missed_rcp += draws && !keep_rcp;
      missed_old += draws && !keep_old;
    }
  }
  std::printf("boxes %ld, holding a drawn pixel %ld, culled %ld; culled but drawn: %ld (rcp %ld; without the "
              "rounding term %ld)\n",
              boxes, drawn_boxes, culled, missed, missed_rcp, missed_old);
  EXPECT(missed == 0 && missed_rcp == 0, "culled boxes hold drawn pixels: %ld / %ld", missed, missed_rcp);
  EXPECT(missed_old > 0, "the cases do not reach the rounding regime (old slack lost nothing)");
  EXPECT(culled > boxes / 10, "the cull drops too little to be useful: %ld of %ld", culled, boxes);
  if (test::failures() == 0) std::printf("test_cull: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
