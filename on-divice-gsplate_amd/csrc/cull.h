// Quadrant cull of the rasterizers: may a splat draw anywhere in a box of
// pixel centres?  Host and device code, so the CPU tests run the same source.
#ifndef GS_CULL_H
#define GS_CULL_H

#include <hip/hip_runtime.h>

#include <cmath>

namespace gsh {

// Smallest q(u, v) = a u^2 / 2 + b u v + c v^2 / 2 (a, c > 0, ac > b^2) over
// the box [u0, u1] x [v0, v1]: 0 when it holds the origin, else the least of
// the four edges' minima (each a clamped 1D vertex).  ia, ic: 1/a, 1/c (an
// approximate reciprocal will do: a vertex only has to be near an edge's
// minimum, and the cull's slack covers it).
__host__ __device__ inline float box_min_quadratic(float a, float b, float c, float ia, float ic, float u0, float u1,
                                                   float v0, float v1) {
  if (u0 <= 0.0f && 0.0f <= u1 && v0 <= 0.0f && 0.0f <= v1) return 0.0f;
  auto q = [&](float u, float v) { return 0.5f * a * u * u + b * u * v + 0.5f * c * v * v; };
  const float va = fminf(fmaxf(-b * u0 * ic, v0), v1), vb = fminf(fmaxf(-b * u1 * ic, v0), v1);
  const float ua = fminf(fmaxf(-b * v0 * ia, u0), u1), ub = fminf(fmaxf(-b * v1 * ia, u0), u1);
  return fminf(fminf(q(u0, va), q(u1, vb)), fminf(q(ua, v0), q(ub, v1)));
}

// False only when no point of the box can reach q <= limit (limit =
// ln(255 o): the splat draws where o exp(-q) >= 1/255) under fp32 evaluation.
// The slack covers (1) the vertex being only near an edge's minimum (q grows
// quadratically away from it: 0.02 + 0.01 |limit|) and (2) rounding: q is a
// sum of terms that cancel, each as large as `scale` (its bound over the
// box), and both this minimum and the rasterizers' per-pixel power round to a
// few ulps of `scale`; 1e-5 is about 170 ulps.  Without (2), needle splats
// (major-axis variance ~1e6 px^2, q quantised to 1/16 far out along the axis)
// lost quadrants that hold pixels they draw.
__host__ __device__ inline bool quadrant_may_draw(float a, float b, float c, float ia, float ic, float limit, float u0,
                                                  float u1, float v0, float v1) {
  const float m = box_min_quadratic(a, b, c, ia, ic, u0, u1, v0, v1);
  const float um = fmaxf(fabsf(u0), fabsf(u1)), vm = fmaxf(fabsf(v0), fabsf(v1));
  const float scale = 0.5f * fabsf(a) * um * um + fabsf(b) * um * vm + 0.5f * fabsf(c) * vm * vm;
  const float slack = 0.02f + 0.01f * fabsf(limit) + 1e-5f * scale;
  return m <= limit + slack;
}

}  // namespace gsh

#endif  // GS_CULL_H
