// Per-slot update: the projection adjoint (2D splat gradient -> position,
// log scale, quaternion, opacity logit, SH DC and rest), the densification
// statistics, and Adam on all of them (engine/shaders/train_step.glsl:62-240,
// train_step_sh.comp:21-44), in two HBM-bound passes over the slots:
//   k_adam_sh  one thread per (slot, group of 5 SH rest coefficients): the
//              group's colour gradient over the step's views, its share of
//              the view-direction part of dL/dposition (written to the
//              scene's dir_grad plane), then Adam on the group's coefficients
//              (coefficient-major planes: 768 B contiguous per wave and
//              coefficient);
//   k_adam     one thread per slot: the rest of the adjoint per view, plus the
//              dir_grad planes, then Adam on the 16 parameters.
// Splitting the SH rest out keeps its 45 floats out of k_adam's registers.
//
// Densification (train_densify_mark.comp, train_densify_spawn.comp) by stream
// compaction: the mark kernel flags prune / clone / split per slot, two
// exclusive scans rank the released slots (the free list) and the spawning
// ones, and the spawn kernel puts the k-th spawner into the k-th free slot or
// past the slots in use.  The slot each new gaussian lands in is therefore a
// function of the flags alone, not of atomic order.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "gs_common.h"
#include "gs_internal.h"
#include "wave.h"

namespace {

using namespace gsd;

struct AdamArgs {
  float4* __restrict__ params;
  float4* __restrict__ m_params;
  float4* __restrict__ v_params;
  float* __restrict__ sh;
  float* __restrict__ m_sh;
  float* __restrict__ v_sh;
  float4* __restrict__ stats;
  float4* __restrict__ dir_grad;           // [kShGroups][capacity]
  const gsh::ViewDev* __restrict__ views;  // the scene's views; views[0].f holds the optimizer's rates
  uint32_t capacity;
  uint32_t sh_coeffs;
  uint32_t count;                               // the scene's views
  const gsh::ViewDev* __restrict__ step_views;  // every view of the step (all scenes)
  uint32_t step_count;
};

__device__ __forceinline__ AdamArgs scene_args(const gsh::SceneDev& sc, const gsh::ViewDev* views, uint32_t count) {
  AdamArgs a;
  a.params = sc.params;
  a.m_params = sc.m_params;
  a.v_params = sc.v_params;
  a.sh = sc.sh;
  a.m_sh = sc.m_sh;
  a.v_sh = sc.v_sh;
  a.stats = sc.stats;
  a.dir_grad = sc.dir_grad;
  a.views = views + sc.first_view;
  a.capacity = sc.capacity;
  a.sh_coeffs = sc.sh_coeffs;
  a.count = sc.views;
  a.step_views = views;
  a.step_count = count;
  return a;
}

// The step leaves the model alone when asked to, or when any view of the step
// found more pairs than its buffers hold (its render was cut short; the host
// runs the step again with larger buffers, trainer.cpp:819-836).
__device__ __forceinline__ bool step_overflowed(const AdamArgs& a) {
  // The step's flag (set by k_work_count, which runs after the pair counts
  // are known): one load instead of a dependent chain over every view.
  if (a.views[0].step_overflow != nullptr) return *gmem(a.views[0].step_overflow) != 0u;
  for (uint32_t v = 0; v < a.step_count; ++v) {
    if (uint64_t(a.step_views[v].totals[0]) > a.step_views[v].pair_capacity) return true;
  }
  return false;
}

__device__ __forceinline__ bool skip_update(const AdamArgs& a) {
  return (a.views[0].f.flags & GS_HIP_FLAG_NO_UPDATE) != 0u || step_overflowed(a);
}

// The gradient side of the optimizer (adjoint, SH gradients, Adam's step)
// is compared with the fp64 oracle within a tolerance, so it uses the
// hardware reciprocal and square root (1 ulp) instead of the correctly
// rounded sequences (about ten instructions each); the projection, whose
// records are bit-exact with the oracle, keeps the exact ones.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ float adam1(float p, float g, float& m, float& v, float lr, const gs_hip_frame& f) {
  m = g * (1.0f - f.beta1) + m * f.beta1;
  v = g * g * (1.0f - f.beta2) + v * f.beta2;
  return p - lr * (m * f.bias1) * fast_rcp(fast_sqrt(v * f.bias2) + f.epsilon);
}

__device__ __forceinline__ float4 adam4(float4 p, float4 g, float4* m_ptr, float4* v_ptr, float4 lr,
                                        const gs_hip_frame& f) {
  float4 m = gload(m_ptr), v = gload(v_ptr);
  p.x = adam1(p.x, g.x, m.x, v.x, lr.x, f);
  p.y = adam1(p.y, g.y, m.y, v.y, lr.y, f);
  p.z = adam1(p.z, g.z, m.z, v.z, lr.z, f);
  p.w = adam1(p.w, g.w, m.w, v.w, lr.w, f);
  gstore(m_ptr, m);
  gstore(v_ptr, v);
  return p;
}

__device__ __forceinline__ void add4(float4& a, float x, float y, float z, float w) {
  a.x += x, a.y += y, a.z += z, a.w += w;
}

// SH rest coefficients per k_adam_sh thread, and the groups of 15.
using gsh::kShGroup;
using gsh::kShGroups;

// basis_{K+1}(direction) of SH rest coefficient K (0..14): the terms of
// sh_basis (gs_common.h), one at a time.
template <int K>
__device__ __forceinline__ float sh_rest_basis(float x, float y, float z) {
#pragma clang fp contract(off)
  const float xx = x * x, yy = y * y, zz = z * z;
  if constexpr (K == 0) return -0.4886025119029199f * y;
  else if constexpr (K == 1) return 0.4886025119029199f * z;
  else if constexpr (K == 2) return -0.4886025119029199f * x;
  else if constexpr (K == 3) return 1.0925484305920792f * x * y;
  else if constexpr (K == 4) return -1.0925484305920792f * y * z;
  else if constexpr (K == 5) return 0.31539156525252005f * (2.0f * zz - xx - yy);
  else if constexpr (K == 6) return -1.0925484305920792f * x * z;
  else if constexpr (K == 7) return 0.5462742152960396f * (xx - yy);
  else if constexpr (K == 8) return -0.5900435899266435f * y * (3.0f * xx - yy);
  else if constexpr (K == 9) return 2.890611442640554f * x * y * z;
  else if constexpr (K == 10) return -0.4570457994644658f * y * (4.0f * zz - xx - yy);
  else if constexpr (K == 11) return 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  else if constexpr (K == 12) return -0.4570457994644658f * x * (4.0f * zz - xx - yy);
  else if constexpr (K == 13) return 1.445305721320277f * z * (xx - yy);
  else return -0.5900435899266435f * x * (xx - 3.0f * yy);
}

// d basis_{K+1} / d(direction) of SH rest coefficient K (0..14).
template <int K>
__device__ __forceinline__ void sh_rest_dbasis(float x, float y, float z, float& gx, float& gy, float& gz) {
  constexpr float C1 = 0.4886025119029199f;
  constexpr float A = 1.0925484305920792f, B = 0.31539156525252005f, C = 0.5462742152960396f;
  constexpr float D = -0.5900435899266435f, E = 2.890611442640554f, F = -0.4570457994644658f;
  constexpr float G = 0.3731763325901154f, H = 1.445305721320277f;
  const float xx = x * x, yy = y * y, zz = z * z;
  if constexpr (K == 0) gx = 0.0f, gy = -C1, gz = 0.0f;
  else if constexpr (K == 1) gx = 0.0f, gy = 0.0f, gz = C1;
  else if constexpr (K == 2) gx = -C1, gy = 0.0f, gz = 0.0f;
  else if constexpr (K == 3) gx = A * y, gy = A * x, gz = 0.0f;
  else if constexpr (K == 4) gx = 0.0f, gy = -A * z, gz = -A * y;
  else if constexpr (K == 5) gx = -2.0f * B * x, gy = -2.0f * B * y, gz = 4.0f * B * z;
  else if constexpr (K == 6) gx = -A * z, gy = 0.0f, gz = -A * x;
  else if constexpr (K == 7) gx = 2.0f * C * x, gy = -2.0f * C * y, gz = 0.0f;
  else if constexpr (K == 8) gx = D * 6.0f * x * y, gy = D * (3.0f * xx - 3.0f * yy), gz = 0.0f;
  else if constexpr (K == 9) gx = E * y * z, gy = E * x * z, gz = E * x * y;
  else if constexpr (K == 10) gx = F * (-2.0f * x * y), gy = F * (4.0f * zz - xx - 3.0f * yy), gz = F * 8.0f * y * z;
  else if constexpr (K == 11)
    gx = G * (-6.0f * x * z), gy = G * (-6.0f * y * z), gz = G * (6.0f * zz - 3.0f * xx - 3.0f * yy);
  else if constexpr (K == 12) gx = F * (4.0f * zz - 3.0f * xx - yy), gy = F * (-2.0f * x * y), gz = F * 8.0f * x * z;
  else if constexpr (K == 13) gx = H * 2.0f * x * z, gy = H * (-2.0f * y * z), gz = H * (xx - yy);
  else gx = D * (3.0f * xx - 3.0f * yy), gy = D * (-6.0f * x * y), gz = 0.0f;
}

// The projection adjoint of one view (train_step.glsl:102-219): 2D splat
// gradient g[9] -> dL/d(position, logit, log scale, quaternion, DC) added to
// the accumulators; the view direction's part of dL/dposition (colour through
// the SH rest) comes from k_adam_sh.  kPose: also the slot's share of the
// camera pose gradient, pose[6] = dL/dd, dL/dw (train_step.glsl:196-201).
template <bool kPose = false>
__device__ __forceinline__ void adjoint_view(const gsh::ViewDev& vd, const float g[9], uint32_t clamped, float4 p0,
                                             float4 p1, float4 q, float4& dp0, float4& dp1, float4& dq, float4& dp3,
                                             float* pose = nullptr) {
  const gs_hip_camera& cam = vd.f.cam;
  const float* V = cam.rotation;
  const float tx = V[0] * p0.x + V[1] * p0.y + V[2] * p0.z + cam.translation[0];
  const float ty = V[3] * p0.x + V[4] * p0.y + V[5] * p0.z + cam.translation[1];
  const float tz = V[6] * p0.x + V[7] * p0.y + V[8] * p0.z + cam.translation[2];
  const float fx = cam.fx, fy = cam.fy;
  const float inv_z = 1.0f / tz;
  const float lim_x = vd.lim_x, lim_y = vd.lim_y;
  const float rx = tx * inv_z, ry = ty * inv_z;
  const bool clamp_x = rx < -lim_x || rx > lim_x, clamp_y = ry < -lim_y || ry > lim_y;
  const float ux = fminf(fmaxf(rx, -lim_x), lim_x), uy = fminf(fmaxf(ry, -lim_y), lim_y);
  const float j00 = fx * inv_z, j02 = -fx * ux * inv_z, j11 = fy * inv_z, j12 = -fy * uy * inv_z;
  float T[2][3];
  for (int c = 0; c < 3; ++c) {
    T[0][c] = j00 * V[c] + j02 * V[6 + c];
    T[1][c] = j11 * V[3 + c] + j12 * V[6 + c];
  }
  const float qlen = fmaxf(fast_sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
  const float iq = fast_rcp(qlen);
  const float qw = q.x * iq, qx = q.y * iq, qy = q.z * iq, qz = q.w * iq;
  float R[3][3];
  R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz), R[0][1] = 2.0f * (qx * qy - qw * qz), R[0][2] = 2.0f * (qx * qz + qw * qy);
  R[1][0] = 2.0f * (qx * qy + qw * qz), R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz), R[1][2] = 2.0f * (qy * qz - qw * qx);
  R[2][0] = 2.0f * (qx * qz - qw * qy), R[2][1] = 2.0f * (qy * qz + qw * qx), R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
  const float s[3] = {__expf(p1.x), __expf(p1.y), __expf(p1.z)};
  float M[3][3], Sig[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[r][c] = R[r][c] * s[c];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Sig[r][c] = M[r][0] * M[c][0] + M[r][1] * M[c][1] + M[r][2] * M[c][2];
  float U[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) U[r][c] = T[r][0] * Sig[0][c] + T[r][1] * Sig[1][c] + T[r][2] * Sig[2][c];
  const float A = U[0][0] * T[0][0] + U[0][1] * T[0][1] + U[0][2] * T[0][2] + kDilation;
  const float B = U[0][0] * T[1][0] + U[0][1] * T[1][1] + U[0][2] * T[1][2];
  const float C = U[1][0] * T[1][0] + U[1][1] * T[1][1] + U[1][2] * T[1][2] + kDilation;
  // The determinant and the gradient through the inverse in fp64: the conic
  // of a needle splat is ill-conditioned (A C / D reaches ~100 on a 1M scene),
  // and in fp32 that alone moved its rotation gradient by ~1e-4 of the
  // largest rotation gradient of the scene.  The slot loop waits on memory,
  // so the fp64 arithmetic costs no time.
  const double Ad = A, Bd = B, Cd = C;
  const double D = Ad * Cd - Bd * Bd;

  // Colour: clamped channels get no gradient; DC.
  const float b0 = 0.28209479177387814f;
  add4(dp3, (clamped & 1u) ? 0.0f : b0 * g[6], (clamped & 2u) ? 0.0f : b0 * g[7], (clamped & 4u) ? 0.0f : b0 * g[8],
       0.0f);

  // Opacity logit.
  const float op = fast_rcp(1.0f + __expf(-p0.w));
  const float dlogit = g[5] * op * (1.0f - op);

  // Conic (C/D, -B/D, A/D) -> dilated covariance.
  const double ga = g[2], gb = g[3], gc = g[4];
  const double inv_d2 = 1.0 / (D * D);
  const float gA = float((-Cd * Cd * ga + Bd * Cd * gb - Bd * Bd * gc) * inv_d2);
  const float gB = float((2.0 * Bd * Cd * ga - (D + 2.0 * Bd * Bd) * gb + 2.0 * Ad * Bd * gc) * inv_d2);
  const float gC = float((-Bd * Bd * ga + Ad * Bd * gb - Ad * Ad * gc) * inv_d2);
  const float G2[2][2] = {{gA, 0.5f * gB}, {0.5f * gB, gC}};
  float G2T[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) G2T[r][c] = G2[r][0] * T[0][c] + G2[r][1] * T[1][c];
  float dSig[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dSig[r][c] = T[0][r] * G2T[0][c] + T[1][r] * G2T[1][c];
  float dT[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c)
      dT[r][c] = 2.0f * (G2T[r][0] * Sig[0][c] + G2T[r][1] * Sig[1][c] + G2T[r][2] * Sig[2][c]);
  const float dJ00 = dT[0][0] * V[0] + dT[0][1] * V[1] + dT[0][2] * V[2];
  const float dJ02 = dT[0][0] * V[6] + dT[0][1] * V[7] + dT[0][2] * V[8];
  const float dJ11 = dT[1][0] * V[3] + dT[1][1] * V[4] + dT[1][2] * V[5];
  const float dJ12 = dT[1][0] * V[6] + dT[1][1] * V[7] + dT[1][2] * V[8];
  const float inv_z2 = inv_z * inv_z;
  float dtx = 0.0f, dty = 0.0f, dtz = -fx * inv_z2 * dJ00 - fy * inv_z2 * dJ11;
  if (clamp_x) {
    dtz += fx * ux * inv_z2 * dJ02;
  } else {
    dtx += -fx * inv_z2 * dJ02;
    dtz += 2.0f * fx * tx * inv_z2 * inv_z * dJ02;
  }
  if (clamp_y) {
    dtz += fy * uy * inv_z2 * dJ12;
  } else {
    dty += -fy * inv_z2 * dJ12;
    dtz += 2.0f * fy * ty * inv_z2 * inv_z * dJ12;
  }
  dtx += g[0] * fx * inv_z;
  dty += g[1] * fy * inv_z;
  dtz += -(g[0] * fx * tx + g[1] * fy * ty) * inv_z2;
  if constexpr (kPose) {
    // A shift d of the camera moves t by d; a turn w moves t by w x t and the
    // rotation in T = J W by [w]x W, whose share of dL/dw is the antisymmetric
    // part of GW = J^T dT W^T.
    float P[3][3];
    for (int c = 0; c < 3; ++c) {
      P[0][c] = j00 * dT[0][c];
      P[1][c] = j11 * dT[1][c];
      P[2][c] = j02 * dT[0][c] + j12 * dT[1][c];
    }
    auto GW = [&](int r, int c) { return P[r][0] * V[c * 3] + P[r][1] * V[c * 3 + 1] + P[r][2] * V[c * 3 + 2]; };
    pose[0] = dtx;
    pose[1] = dty;
    pose[2] = dtz;
    pose[3] = ty * dtz - tz * dty + (GW(2, 1) - GW(1, 2));
    pose[4] = tz * dtx - tx * dtz + (GW(0, 2) - GW(2, 0));
    pose[5] = tx * dty - ty * dtx + (GW(1, 0) - GW(0, 1));
  }
  add4(dp0, V[0] * dtx + V[3] * dty + V[6] * dtz, V[1] * dtx + V[4] * dty + V[7] * dtz,
       V[2] * dtx + V[5] * dty + V[8] * dtz, dlogit);

  float dM[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      dM[r][c] = 2.0f * (dSig[r][0] * M[0][c] + dSig[r][1] * M[1][c] + dSig[r][2] * M[2][c]);
  add4(dp1, (dM[0][0] * R[0][0] + dM[1][0] * R[1][0] + dM[2][0] * R[2][0]) * s[0],
       (dM[0][1] * R[0][1] + dM[1][1] * R[1][1] + dM[2][1] * R[2][1]) * s[1],
       (dM[0][2] * R[0][2] + dM[1][2] * R[1][2] + dM[2][2] * R[2][2]) * s[2], 0.0f);
  float gR[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) gR[r][c] = dM[r][c] * s[c];
  const float dw = 2.0f * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] +
                           qx * gR[2][1]);
  const float dx = 2.0f * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2.0f * qx * gR[1][1] - qw * gR[1][2] +
                           qz * gR[2][0] + qw * gR[2][1] - 2.0f * qx * gR[2][2]);
  const float dy = 2.0f * (-2.0f * qy * gR[0][0] + qx * gR[0][1] + qw * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] -
                           qw * gR[2][0] + qz * gR[2][1] - 2.0f * qy * gR[2][2]);
  const float dz = 2.0f * (-2.0f * qz * gR[0][0] - qw * gR[0][1] + qx * gR[0][2] + qw * gR[1][0] -
                           2.0f * qz * gR[1][1] + qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
  const float qd = qw * dw + qx * dx + qy * dy + qz * dz;
  add4(dq, (dw - qw * qd) * iq, (dx - qx * qd) * iq, (dy - qy * qd) * iq, (dz - qz * qd) * iq);
}

// Adjoint of every view in `a.views`, summed and averaged, then one Adam step
// on the 16 parameters; densify statistics count every view in which the
// gaussian was visible.  The view-direction part of dL/dposition was summed
// over the views by k_adam_sh (one plane per group of SH rest coefficients).
// Four waves per SIMD: its fp64 determinant would otherwise take it to 132
// VGPRs and three.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_adam(const gsh::SceneDev* __restrict__ scenes, const gsh::ViewDev* __restrict__ views, uint32_t count) {
  const AdamArgs a = scene_args(scenes[blockIdx.y], views, count);
  const gs_hip_frame& f = a.views[0].f;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.slots || skip_update(a)) return;
  const int coeffs = min(sh_rest_coeffs(f.sh_degree), int(a.sh_coeffs));
  float4 p0 = gload(a.params + i * 4u), p1 = gload(a.params + i * 4u + 1u), q = gload(a.params + i * 4u + 2u),
         p3 = gload(a.params + i * 4u + 3u);
  if (p1.w < 0.5f) return;  // free slot
  float4 dp0 = make_float4(0, 0, 0, 0), dp1 = dp0, dq = dp0, dp3 = dp0;
  for (int k0 = 0; k0 < coeffs; k0 += kShGroup) {
    const float4 d = gload(a.dir_grad + size_t(k0 / kShGroup) * a.capacity + i);
    add4(dp0, d.x, d.y, d.z, 0.0f);
  }
  bool drew = false, seen = false;
  float4 st = gload(a.stats + i);
  for (uint32_t v = 0; v < a.count; ++v) {
    const gsh::ViewDev& view = a.views[v];
    if (gmem(view.tiles)[i] == 0u) continue;
    seen = true;
    float g[9];
    const GS_GLOBAL float* g2 = gmem(view.grad2d);
    for (int c = 0; c < 9; ++c) {
      g[c] = g2[size_t(i) * 9u + c];
      drew = drew || g[c] != 0.0f;
    }
    // Densification statistics: view-space positional gradient in NDC units.
    const float gx = g[0] * (0.5f * float(view.f.cam.width)), gy = g[1] * (0.5f * float(view.f.cam.height));
    const uint32_t misc = gmem(view.misc)[i];
    st.x += sqrtf(gx * gx + gy * gy);
    st.y += 1.0f;
    st.z = fmaxf(st.z, float(misc >> 3));
    adjoint_view(view, g, misc & 7u, p0, p1, q, dp0, dp1, dq, dp3);
  }
  if (seen) gstore(a.stats + i, st);
  if (!drew && (f.flags & GS_HIP_FLAG_SPARSE_ADAM) != 0u) return;
  const float inv = 1.0f / float(a.count);  // the mean over the views
  const float4 lr0 = make_float4(f.lr_position, f.lr_position, f.lr_position, f.lr_opacity);
  const float4 lr1 = make_float4(f.lr_scale, f.lr_scale, f.lr_scale, 0.0f);
  const float4 lrq = make_float4(f.lr_rotation, f.lr_rotation, f.lr_rotation, f.lr_rotation);
  const float4 lr3 = make_float4(f.lr_sh_dc, f.lr_sh_dc, f.lr_sh_dc, 0.0f);
  auto scale4 = [inv](float4 v) { return make_float4(v.x * inv, v.y * inv, v.z * inv, v.w * inv); };
  gstore(a.params + i * 4u, adam4(p0, scale4(dp0), a.m_params + i * 4u, a.v_params + i * 4u, lr0, f));
  gstore(a.params + i * 4u + 1u, adam4(p1, scale4(dp1), a.m_params + i * 4u + 1u, a.v_params + i * 4u + 1u, lr1, f));
  gstore(a.params + i * 4u + 2u, adam4(q, scale4(dq), a.m_params + i * 4u + 2u, a.v_params + i * 4u + 2u, lrq, f));
  gstore(a.params + i * 4u + 3u, adam4(p3, scale4(dp3), a.m_params + i * 4u + 3u, a.v_params + i * 4u + 3u, lr3, f));
}

// Camera pose gradient of each view that asks for it (GS_HIP_FLAG_POSE_GRAD,
// train_step.glsl:196-201, 253-269): every visible slot's share from the
// adjoint's camera-space terms, summed per block of 256 slots into the view's
// pose_grad[block] = (dL/dd xyz, 0), (dL/dw xyz, 0); the host adds the blocks
// in double.  Runs before k_adam, on the parameters the step rendered with:
// 64 B of parameters and 36 B of GRAD2D per visible slot and view.  A step
// whose pairs overflowed writes zeros (its gradients are incomplete; it runs
// again).  Each view's gradient is that of its own loss.
__global__ void __launch_bounds__(256) k_pose_grad(const gsh::SceneDev* __restrict__ scenes,
                                                   const gsh::ViewDev* __restrict__ views, uint32_t count) {
  const AdamArgs a = scene_args(scenes[blockIdx.y], views, count);
  if (blockIdx.x * 256u >= a.capacity) return;  // past this scene's slots: the whole block
  const bool overflow = step_overflowed(a);
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  bool live = !overflow && i < a.views[0].f.slots;
  float4 p0 = make_float4(0, 0, 0, 0), p1 = p0, q = p0;
  if (live) {
    p0 = a.params[i * 4u];
    p1 = a.params[i * 4u + 1u];
    q = a.params[i * 4u + 2u];
    live = p1.w >= 0.5f;  // not a free slot
  }
  __shared__ float s_part[6][4];
  const int wave = threadIdx.x >> 6;
  for (uint32_t v = 0; v < a.count; ++v) {
    const gsh::ViewDev& view = a.views[v];
    if ((view.f.flags & GS_HIP_FLAG_POSE_GRAD) == 0u) continue;  // the same for the whole block
    float pose[6] = {0, 0, 0, 0, 0, 0};
    if (live && view.tiles[i] != 0u) {
      float g[9];
      for (int c = 0; c < 9; ++c) g[c] = gmem(view.grad2d)[size_t(i) * 9u + c];
      float4 dp0 = make_float4(0, 0, 0, 0), dp1 = dp0, dq = dp0, dp3 = dp0;
      adjoint_view<true>(view, g, 0u, p0, p1, q, dp0, dp1, dq, dp3, pose);
    }
    for (int k = 0; k < 6; ++k) {
      const float w = wave_sum(pose[k]);
      if ((threadIdx.x & 63) == 0) s_part[k][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      float r[3];
      for (int k = 0; k < 3; ++k) {
        const int c = int(threadIdx.x) * 3 + k;
        r[k] = s_part[c][0] + s_part[c][1] + s_part[c][2] + s_part[c][3];
      }
      view.pose_grad[size_t(blockIdx.x) * 2u + threadIdx.x] = make_float4(r[0], r[1], r[2], 0.0f);
    }
    __syncthreads();
  }
}

// One group of kShGroup SH rest coefficients of one slot (train_step.glsl:
// 129-155, train_step_sh.comp:21-44): for every view of the step that sees
// the slot, the colour gradient dcol (clamped channels 0) against the basis
// at the view direction, and the group's share of d(colour)/d(direction)
// against dcol, carried to dL/dposition through the direction's
// normalisation; the latter summed over the views into dir_grad[group] for
// k_adam, then Adam on the group's coefficients at the views' mean gradient.
// f(integral_constant<int, D>) for D in [D0, N).
template <int D0, int N, class F>
__device__ __forceinline__ void unroll_const(F& f) {
  if constexpr (D0 < N) {
    f(std::integral_constant<int, D0>{});
    unroll_const<D0 + 1, N>(f);
  }
}

template <int kGroup>
__device__ __forceinline__ void sh_group(const AdamArgs& a, const gs_hip_frame& f, uint32_t i, int coeffs) {
  constexpr int k0 = kGroup * kShGroup;
  const float4 p0 = gload(a.params + i * 4u);
  float rest[kShGroup][3], g[kShGroup][3];
#pragma unroll
  for (int d = 0; d < kShGroup; ++d) {
    g[d][0] = g[d][1] = g[d][2] = 0.0f;
    rest[d][0] = rest[d][1] = rest[d][2] = 0.0f;
    if (k0 + d < coeffs) {
      const GS_GLOBAL float* sk = gmem(a.sh) + (size_t(k0 + d) * a.capacity + i) * 3u;
      rest[d][0] = sk[0];
      rest[d][1] = sk[1];
      rest[d][2] = sk[2];
    }
  }
  const bool sparse = (f.flags & GS_HIP_FLAG_SPARSE_ADAM) != 0u;
  bool drew = false;
  float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
  // Each view's colour gradient and clamp bits, requested two views ahead of
  // their use (GRAD2D is zero where the view does not see the slot: the
  // projection clears it, so no tile-count test is needed).
  auto fetch = [&](uint32_t v) {
    if (v >= a.count) return make_float4(0, 0, 0, 0);
    const gsh::ViewDev& view = a.views[v];
    const GS_GLOBAL float* col = gmem(view.grad2d) + size_t(i) * 9u + 6u;
    return make_float4(col[0], col[1], col[2], __uint_as_float(gmem(view.misc)[i]));
  };
  float4 ahead0 = fetch(0), ahead1 = fetch(1);
#pragma unroll
  for (uint32_t v = 0; v < GS_HIP_MAX_VIEWS; ++v) {
    if (v >= a.count) break;
    const gsh::ViewDev& view = a.views[v];
    const float4 gc = ahead0;
    ahead0 = ahead1;
    ahead1 = fetch(v + 2);
    drew = drew || gc.x != 0.0f || gc.y != 0.0f || gc.z != 0.0f;
    if (sparse && !drew) {
      for (int c = 0; c < 6; ++c) drew = drew || gmem(view.grad2d)[size_t(i) * 9u + c] != 0.0f;
    }
    const uint32_t clamped = __float_as_uint(gc.w) & 7u;
    const float dcol[3] = {clamped & 1u ? 0.0f : gc.x, clamped & 2u ? 0.0f : gc.y, clamped & 4u ? 0.0f : gc.z};
    if (dcol[0] == 0.0f && dcol[1] == 0.0f && dcol[2] == 0.0f) continue;
    const float drx = p0.x - view.f.cam.position[0], dry = p0.y - view.f.cam.position[1];
    const float drz = p0.z - view.f.cam.position[2];
    const float dlen = fmaxf(fast_sqrt(drx * drx + dry * dry + drz * drz), 1e-12f);
    const float il = fast_rcp(dlen);
    const float x = drx * il, y = dry * il, z = drz * il;
    float ddx = 0.0f, ddy = 0.0f, ddz = 0.0f;
    auto term = [&](auto kc) {
      constexpr int d = decltype(kc)::value;
      if (k0 + d >= coeffs) return;
      const float b = sh_rest_basis<k0 + d>(x, y, z);
      for (int c = 0; c < 3; ++c) g[d][c] += b * dcol[c];
      const float w = rest[d][0] * dcol[0] + rest[d][1] * dcol[1] + rest[d][2] * dcol[2];
      float gx, gy, gz;
      sh_rest_dbasis<k0 + d>(x, y, z, gx, gy, gz);
      ddx += w * gx;
      ddy += w * gy;
      ddz += w * gz;
    };
    unroll_const<0, kShGroup>(term);
    const float dd = x * ddx + y * ddy + z * ddz;
    dpx += (ddx - x * dd) * il;
    dpy += (ddy - y * dd) * il;
    dpz += (ddz - z * dd) * il;
  }
  gstore(a.dir_grad + size_t(kGroup) * a.capacity + i, make_float4(dpx, dpy, dpz, 0.0f));
  if (sparse && !drew) return;  // kept by sparse Adam
  const float inv = 1.0f / float(a.count);
  // A slot's three channels of a coefficient are stored together per plane,
  // so each plane's three stores merge into one 12-byte store instruction.
  auto plane = [&](float* base, int d) { return gmem(base) + (size_t(k0 + d) * a.capacity + i) * 3u; };
  float m[kShGroup][3], vv[kShGroup][3];
  for (int d = 0; d < kShGroup; ++d) {
    if (k0 + d >= coeffs) break;
    for (int c = 0; c < 3; ++c) m[d][c] = plane(a.m_sh, d)[c], vv[d][c] = plane(a.v_sh, d)[c];
  }
  for (int d = 0; d < kShGroup; ++d) {
    if (k0 + d >= coeffs) break;
    float p[3];
    for (int c = 0; c < 3; ++c) p[c] = adam1(rest[d][c], g[d][c] * inv, m[d][c], vv[d][c], f.lr_sh_rest, f);
    GS_GLOBAL float* ps = plane(a.sh, d);
    GS_GLOBAL float* pm = plane(a.m_sh, d);
    GS_GLOBAL float* pv = plane(a.v_sh, d);
    for (int c = 0; c < 3; ++c) ps[c] = p[c];
    for (int c = 0; c < 3; ++c) pm[c] = m[d][c];
    for (int c = 0; c < 3; ++c) pv[c] = vv[d][c];
  }
}

// A block is the groups of 64 slots, one wave each: every wave reads its
// coefficient planes coalesced, and the views' per-slot gradients, clamp
// bits and tile counts that all of them read come from the CU's cache after
// the first.
constexpr int kShBlock = 64 * kShGroups;

template <int G = 0>
__device__ __forceinline__ void sh_group_of(int group, const AdamArgs& a, const gs_hip_frame& f, uint32_t i, int coeffs) {
  if constexpr (G < kShGroups) {
    if (group == G) {
      sh_group<G>(a, f, i, coeffs);
    } else {
      sh_group_of<G + 1>(group, a, f, i, coeffs);
    }
  }
}

__global__ void __launch_bounds__(kShBlock) k_adam_sh(const gsh::SceneDev* __restrict__ scenes,
                                                      const gsh::ViewDev* __restrict__ views, uint32_t count) {
  const AdamArgs a = scene_args(scenes[blockIdx.y], views, count);
  const gs_hip_frame& f = a.views[0].f;
  const int group = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const uint32_t i = blockIdx.x * 64u + (threadIdx.x & 63u);
  const int coeffs = min(sh_rest_coeffs(f.sh_degree), int(a.sh_coeffs));
  if (group * kShGroup >= coeffs || i >= f.slots || skip_update(a)) return;
  if (gload(a.params + i * 4u + 1u).w < 0.5f) return;  // free slot
  sh_group_of(group, a, f, i, coeffs);
}

// ---- densification ----------------------------------------------------------
struct DensifyArgs {
  float4* __restrict__ params;
  float4* __restrict__ m_params;
  float4* __restrict__ v_params;
  float* __restrict__ sh;
  float* __restrict__ m_sh;
  float* __restrict__ v_sh;
  float4* __restrict__ stats;
  uint32_t* __restrict__ action;    // 0 none, 1 clone, 2 split
  uint32_t* __restrict__ release;   // 1: slot goes on the free list
  uint32_t* __restrict__ spawn;     // 1: slot spawns one gaussian
  uint32_t* __restrict__ free_rank;
  uint32_t* __restrict__ spawn_rank;
  uint32_t* __restrict__ free_list;
  uint32_t* __restrict__ totals;    // [5] pruned, [6] free count, [7] spawn count
  uint32_t capacity;
  uint32_t sh_coeffs;
};

__device__ __forceinline__ void clear_moments(const DensifyArgs& a, uint32_t slot) {
  for (int k = 0; k < 4; ++k) {
    a.m_params[slot * 4u + k] = make_float4(0, 0, 0, 0);
    a.v_params[slot * 4u + k] = make_float4(0, 0, 0, 0);
  }
  for (uint32_t k = 0; k < a.sh_coeffs; ++k) {
    const size_t o = (size_t(k) * a.capacity + slot) * 3u;
    for (int c = 0; c < 3; ++c) a.m_sh[o + c] = 0.0f, a.v_sh[o + c] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_densify_mark(gs_hip_frame f, DensifyArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.slots) return;
  const float4 p0 = a.params[i * 4u], p1 = a.params[i * 4u + 1u];
  uint32_t action = 0u;
  bool release = p1.w < 0.5f;
  if (!release) {
    const float4 st = a.stats[i];
    // Decisions on the portable exp, so the CPU oracle reproduces the mask bit for bit.
    const float max_scale = gs_expf(fmaxf(p1.x, fmaxf(p1.y, p1.z)));
    bool prune = 1.0f / (1.0f + gs_expf(-p0.w)) < f.prune_opacity;
    if (f.prune_max_scale > 0.0f && max_scale > f.prune_max_scale) prune = true;
    if (f.prune_max_radius > 0.0f && st.z > f.prune_max_radius) prune = true;
    if (prune) {
      release = true;
      atomicAdd(&a.totals[5], 1u);
      a.params[i * 4u] = make_float4(0.0f, 0.0f, 0.0f, -20.0f);
      a.params[i * 4u + 1u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      a.params[i * 4u + 2u] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
      a.params[i * 4u + 3u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      clear_moments(a, i);
      for (uint32_t k = 0; k < a.sh_coeffs; ++k) {
        const size_t o = (size_t(k) * a.capacity + i) * 3u;
        a.sh[o] = a.sh[o + 1] = a.sh[o + 2] = 0.0f;
      }
    } else if (st.y > 0.0f && st.x / st.y >= f.grad_threshold) {
      action = max_scale <= f.clone_max_scale ? 1u : 2u;
    }
  }
  a.action[i] = action;
  a.release[i] = release ? 1u : 0u;
  a.spawn[i] = action != 0u ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_free_list(uint32_t slots, DensifyArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < slots && a.release[i] != 0u) a.free_list[a.free_rank[i]] = i;
}

// PCG hash (O'Neill's permuted congruential generator, RXS-M-XS output) and
// Box-Muller normals, as the reference samples split halves
// (train_common.glsl:144-161).
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

__device__ __forceinline__ float uniform01(uint32_t& state) {
  state = pcg_hash(state);
  return float(state >> 8u) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ void normal3(uint32_t& state, float n[3]) {
  const float u1 = fmaxf(uniform01(state), 1e-7f), u2 = uniform01(state);
  const float u3 = fmaxf(uniform01(state), 1e-7f), u4 = uniform01(state);
  const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
  n[0] = r1 * cosf(6.283185307f * u2);
  n[1] = r1 * sinf(6.283185307f * u2);
  n[2] = r2 * cosf(6.283185307f * u4);
}

__global__ void __launch_bounds__(256) k_densify_spawn(gs_hip_frame f, DensifyArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.slots) return;
  const uint32_t action = a.action[i];
  if (action == 0u) return;
  const uint32_t k = a.spawn_rank[i];
  const uint32_t free_count = a.totals[6];
  const uint32_t target = k < free_count ? a.free_list[k] : f.slots + (k - free_count);
  if (target >= a.capacity) return;
  const float4 p0 = a.params[i * 4u], p1 = a.params[i * 4u + 1u], q = a.params[i * 4u + 2u],
               p3 = a.params[i * 4u + 3u];
  for (uint32_t kk = 0; kk < a.sh_coeffs; ++kk) {
    const size_t src = (size_t(kk) * a.capacity + i) * 3u, dst = (size_t(kk) * a.capacity + target) * 3u;
    for (int c = 0; c < 3; ++c) a.sh[dst + c] = a.sh[src + c];
  }
  clear_moments(a, target);
  if (action == 1u) {
    a.params[target * 4u] = p0;
    a.params[target * 4u + 1u] = p1;
  } else {
    uint32_t rng = pcg_hash(i ^ pcg_hash(f.seed));
    const float s[3] = {expf(p1.x), expf(p1.y), expf(p1.z)};
    const float ql = fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
    const float w = q.x / ql, x = q.y / ql, y = q.z / ql, z = q.w / ql;
    const float R[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
                           {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
                           {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
    float na[3], nb[3];
    normal3(rng, na);
    normal3(rng, nb);
    float pa[3], pb[3];
    for (int r = 0; r < 3; ++r) {
      pa[r] = (&p0.x)[r] + R[r][0] * na[0] * s[0] + R[r][1] * na[1] * s[1] + R[r][2] * na[2] * s[2];
      pb[r] = (&p0.x)[r] + R[r][0] * nb[0] * s[0] + R[r][1] * nb[1] * s[1] + R[r][2] * nb[2] * s[2];
    }
    const float shrink = logf(f.split_divisor);
    const float4 ls = make_float4(p1.x - shrink, p1.y - shrink, p1.z - shrink, 1.0f);
    a.params[i * 4u] = make_float4(pa[0], pa[1], pa[2], p0.w);
    a.params[i * 4u + 1u] = ls;
    a.params[target * 4u] = make_float4(pb[0], pb[1], pb[2], p0.w);
    a.params[target * 4u + 1u] = ls;
    clear_moments(a, i);
  }
  a.params[target * 4u + 2u] = q;
  a.params[target * 4u + 3u] = p3;
}

__global__ void __launch_bounds__(256) k_reset_opacity(uint32_t slots, float max_logit, float4* __restrict__ params,
                                                       float4* __restrict__ m_params, float4* __restrict__ v_params) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots || params[i * 4u + 1u].w < 0.5f) return;
  params[i * 4u].w = fminf(params[i * 4u].w, max_logit);
  m_params[i * 4u].w = 0.0f;
  v_params[i * 4u].w = 0.0f;
}

__global__ void __launch_bounds__(256) k_zero_stats(float4* __restrict__ stats, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) stats[i] = make_float4(0, 0, 0, 0);
}

__global__ void k_zero_counters(uint32_t* __restrict__ totals) {
  if (threadIdx.x < 3) totals[5 + threadIdx.x] = 0u;
}

}  // namespace

namespace gsh {

int launch_adam(gs_hip_workspace* ws, uint32_t count) {
  if (count == 0 || count != ws->shape.views) return set_error(GS_HIP_E_ARGUMENT, "adam: view count out of range");
  if (ws->shape.max_capacity == 0) return 0;
  const dim3 grid(div_up(ws->shape.max_capacity, 256), ws->shape.scenes);
  if (ws->shape.any_sh) {  // first: it reads the positions k_adam updates
    timer_begin(ws, "adam_sh");
    k_adam_sh<<<dim3(div_up(ws->shape.max_capacity, 64), grid.y), kShBlock, 0, ws->stream>>>(ws->dev_scenes,
                                                                                          ws->dev_views, count);
    GS_CHECK_LAUNCH("k_adam_sh");
    timer_end(ws);
  }
  timer_begin(ws, "adam");
  k_adam<<<grid, 256, 0, ws->stream>>>(ws->dev_scenes, ws->dev_views, count);
  GS_CHECK_LAUNCH("k_adam");
  timer_end(ws);
  return 0;
}

int launch_pose(gs_hip_workspace* ws, uint32_t count) {
  if (ws->shape.max_capacity == 0) return 0;
  const dim3 grid(div_up(ws->shape.max_capacity, 256), ws->shape.scenes);
  timer_begin(ws, "pose");
  k_pose_grad<<<grid, 256, 0, ws->stream>>>(ws->dev_scenes, ws->dev_views, count);
  GS_CHECK_LAUNCH("k_pose_grad");
  timer_end(ws);
  return 0;
}

namespace {

// What a captured graph was recorded for.
uint64_t step_signature(uint32_t count, bool depth16, bool train, bool losses) {
  return uint64_t(count) | (uint64_t(depth16) << 8) | (uint64_t(train) << 9) | (uint64_t(losses) << 10);
}

// The stages of one batched step over views [0, count), in stream order.
int launch_stages(gs_hip_workspace* ws, uint32_t count, bool depth16, bool train, bool losses) {
  int rc = launch_project(ws, count);
  if (rc == 0) rc = launch_depth_sort_and_scan(ws, count, depth16);
  if (rc == 0) rc = launch_emit_and_tile_sort(ws, count);
  if (rc == 0) rc = launch_forward(ws, count);
  if (rc == 0 && (train || losses)) rc = launch_loss(ws, count);
  if (rc == 0 && train) rc = launch_backward(ws, count);
  if (rc == 0 && train && ws->shape.any_pose) rc = launch_pose(ws, count);
  if (rc == 0 && train) rc = launch_adam(ws, count);
  if (rc == 0) rc = launch_readback(ws, count, losses);
  return rc;
}

// Replays the captured step (capturing it first when the views, buffers,
// launch shapes or step kind changed: the table signature covers the buffers).
int launch_graph(gs_hip_workspace* ws, uint32_t count, bool depth16, bool train, bool losses) {
  const uint64_t sig = step_signature(count, depth16, train, losses) ^ (ws->table_signature * 0x100000001b3ull);
  if (ws->graph_exec == nullptr || ws->graph_signature != sig) {
    drop_graph(ws);
    GS_HIP_TRY(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_stages(ws, count, depth16, train, losses);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(ws->stream, &graph);
    if (rc != 0) {
      if (graph != nullptr) hipGraphDestroy(graph);
      return rc;
    }
    if (end != hipSuccess) return hip_error(end, "hipStreamEndCapture");
    ws->graph = graph;
    GS_HIP_TRY(hipGraphInstantiate(&ws->graph_exec, ws->graph, nullptr, nullptr, 0));
    ws->graph_signature = sig;
  }
  GS_HIP_TRY(hipGraphLaunch(ws->graph_exec, ws->stream));
  return 0;
}

}  // namespace

int launch_step(gs_hip_workspace* ws, uint32_t count, bool depth16, bool train, bool losses) {
  const bool graph = ws->tuning.graph != 0u && !ws->timer.enabled;
  return graph ? launch_graph(ws, count, depth16, train, losses) : launch_stages(ws, count, depth16, train, losses);
}

namespace {

int check_frames(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count) {
  if (count == 0 || count > ws->views.size()) {
    return set_error(GS_HIP_E_ARGUMENT, "step: view count exceeds the workspace's views");
  }
  for (uint32_t v = 0; v < count; ++v) {
    const gs_hip_frame& f = frames[v];
    if (f.slots > ws->capacity || f.slots != frames[0].slots) {
      return set_error(GS_HIP_E_ARGUMENT, "step: slots differ between views or exceed capacity");
    }
    if (f.cam.width != ws->width || f.cam.height != ws->height) {
      return set_error(GS_HIP_E_ARGUMENT, "step: camera size differs from the workspace");
    }
    if (f.tiles_x != ws->tiles_x || f.tiles_y != ws->tiles_y) {
      return set_error(GS_HIP_E_ARGUMENT, "step: tile grid differs from the workspace");
    }
  }
  return 0;
}

}  // namespace

// One batched step: every stage of every view, pair counts on the device;
// when a view's pairs overflow the buffers the optimizer skipped the update,
// so grow the buffers and run it again (trainer.cpp:812-836).
int run_step(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets, bool train,
             bool losses, uint64_t* pairs, double* loss_sums) {
  int rc = check_frames(ws, frames, count);
  if (rc != 0) return rc;
  const bool depth16 = (frames[0].flags & GS_HIP_FLAG_DEPTH16) != 0u;
  // The caller's targets are read in place (through the view table), not copied.
  // Headroom over the largest count seen (trainer.cpp:588-592).
  const uint64_t wanted = ws->max_pairs_seen + ws->max_pairs_seen / 8;
  if (wanted > view0(ws).pair_capacity) {
    rc = reserve_pairs(ws, wanted);
    if (rc != 0) return rc;
  }
  for (int attempt = 0;; ++attempt) {
    rc = upload_table(ws, frames, count, targets);
    if (rc != 0) return rc;
    rc = launch_step(ws, count, depth16, train, losses);
    if (rc != 0) return rc;
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
    uint64_t most = 0;
    for (uint32_t v = 0; v < count; ++v) {
      const uint64_t p = ws->host_totals[v * 16];
      ws->views[v]->pairs = p;
      most = std::max(most, p);
    }
    ws->max_pairs_seen = std::max(ws->max_pairs_seen, most);
    if (most <= view0(ws).pair_capacity) break;
    ++ws->overflow_steps;
    if (attempt > 4) return set_error(GS_HIP_E_CAPACITY, "step: pair buffers keep overflowing");
    rc = reserve_pairs(ws, most);
    if (rc != 0) return rc;
  }
  for (uint32_t v = 0; v < count; ++v) {
    if (pairs != nullptr) pairs[v] = ws->views[v]->pairs;
    if (losses && loss_sums != nullptr) sum_losses(ws, v, loss_sums + 3 * v);
  }
  return 0;
}

}  // namespace gsh

extern "C" {

int gs_hip_adam(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "adam: null argument");
  int rc = gsh::upload_table(ws, f, 1);
  if (rc == 0 && ws->shape.any_pose) rc = gsh::launch_pose(ws, 1);
  return rc != 0 ? rc : gsh::launch_adam(ws, 1);
}

int gs_hip_pose_gradient(gs_hip_workspace* ws, uint32_t view, double grad[6]) {
  if (ws == nullptr || grad == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "pose_gradient: null argument");
  if (view >= ws->views.size()) return gsh::set_error(GS_HIP_E_ARGUMENT, "pose_gradient: no such view");
  const gsh::View& v = *ws->views[view];
  const uint32_t blocks = gsh::div_up(std::max<uint32_t>(ws->capacity, 1), 256);
  std::vector<float> part(size_t(blocks) * 8);
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  GS_HIP_TRY(hipMemcpy(part.data(), v.pose_grad, part.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int c = 0; c < 6; ++c) grad[c] = 0.0;
  for (uint32_t b = 0; b < blocks; ++b) {
    for (int c = 0; c < 3; ++c) {
      grad[c] += part[size_t(b) * 8 + c];
      grad[3 + c] += part[size_t(b) * 8 + 4 + c];
    }
  }
  return 0;
}

int gs_hip_densify(gs_hip_workspace* ws, const gs_hip_frame* f, uint32_t* slots, uint32_t counts[3]) {
  if (ws == nullptr || f == nullptr || slots == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "densify: null");
  if (f->slots != *slots) return gsh::set_error(GS_HIP_E_ARGUMENT, "densify: frame slots differ from *slots");
  if (counts != nullptr) counts[0] = counts[1] = counts[2] = 0;
  const uint32_t n = *slots;
  if (n == 0) return 0;
  gsh::View& v0 = gsh::view0(ws);
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  DensifyArgs a;
  a.params = reinterpret_cast<float4*>(ws->params);
  a.m_params = reinterpret_cast<float4*>(ws->m_params);
  a.v_params = reinterpret_cast<float4*>(ws->v_params);
  a.sh = ws->sh;
  a.m_sh = ws->m_sh;
  a.v_sh = ws->v_sh;
  a.stats = reinterpret_cast<float4*>(ws->stats);
  // View 0's per-step scratch: the next projection rewrites all of it.
  a.action = v0.flags;
  a.release = v0.depth_keys;
  a.spawn = v0.depth_order;
  a.free_rank = v0.offsets;
  a.spawn_rank = v0.scan_tmp;
  a.free_list = v0.tiles;
  a.totals = v0.totals;
  a.capacity = ws->capacity;
  a.sh_coeffs = ws->sh_coeffs;
  const uint32_t blocks = gsh::div_up(n, 256);
  gsh::timer_begin(ws, "densify");
  k_zero_counters<<<1, 64, 0, ws->stream>>>(v0.totals);
  GS_CHECK_LAUNCH("k_zero_counters");
  k_densify_mark<<<blocks, 256, 0, ws->stream>>>(*f, a);
  GS_CHECK_LAUNCH("k_densify_mark");
  int rc = gsh::exclusive_scan(ws, v0, a.release, nullptr, a.free_rank, n, v0.totals + 6);
  if (rc != 0) return rc;
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));  // the scan segment is reused below
  rc = gsh::exclusive_scan(ws, v0, a.spawn, nullptr, a.spawn_rank, n, v0.totals + 7);
  if (rc != 0) return rc;
  k_free_list<<<blocks, 256, 0, ws->stream>>>(n, a);
  GS_CHECK_LAUNCH("k_free_list");
  k_densify_spawn<<<blocks, 256, 0, ws->stream>>>(*f, a);
  GS_CHECK_LAUNCH("k_densify_spawn");
  k_zero_stats<<<gsh::div_up(ws->capacity, 256), 256, 0, ws->stream>>>(a.stats, ws->capacity);
  GS_CHECK_LAUNCH("k_zero_stats");
  GS_HIP_TRY(hipMemcpyAsync(ws->host_totals, v0.totals, 32, hipMemcpyDeviceToHost, ws->stream));
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  gsh::timer_end(ws);
  const uint32_t pruned = ws->host_totals[5], free_slots = ws->host_totals[6], spawned = ws->host_totals[7];
  const uint32_t room = free_slots + (ws->capacity - n);
  const uint32_t placed = spawned < room ? spawned : room;
  if (placed > free_slots) *slots = n + (placed - free_slots);
  if (counts != nullptr) {
    counts[0] = free_slots;
    counts[1] = placed;
    counts[2] = pruned;
  }
  return 0;
}

int gs_hip_reset_opacity(gs_hip_workspace* ws, const gs_hip_frame* f, float max_logit) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "reset_opacity: null argument");
  if (f->slots == 0) return 0;
  k_reset_opacity<<<gsh::div_up(f->slots, 256), 256, 0, ws->stream>>>(
      f->slots, max_logit, reinterpret_cast<float4*>(ws->params), reinterpret_cast<float4*>(ws->m_params),
      reinterpret_cast<float4*>(ws->v_params));
  GS_CHECK_LAUNCH("k_reset_opacity");
  return 0;
}

int gs_hip_train_step(gs_hip_workspace* ws, const gs_hip_frame* f, uint64_t* pairs, double losses[3]) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "train_step: null argument");
  return gsh::run_step(ws, f, 1, nullptr, true, losses != nullptr, pairs, losses);
}

int gs_hip_train_views(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count,
                       const void* const* targets, uint64_t* pairs, double* losses) {
  if (ws == nullptr || frames == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "train_views: null argument");
  return gsh::run_step(ws, frames, count, targets, true, losses != nullptr, pairs, losses);
}

int gs_hip_render(gs_hip_workspace* ws, const gs_hip_frame* f, uint8_t* rgba8_out) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "render: null argument");
  const int rc = gsh::run_step(ws, f, 1, nullptr, false, false, nullptr, nullptr);
  if (rc != 0) return rc;
  if (rgba8_out != nullptr) {
    gsh::View& v = gsh::view0(ws);
    const size_t pixels = size_t(ws->width) * ws->height;
    std::vector<float> im(pixels * 4);
    GS_HIP_TRY(hipMemcpyAsync(im.data(), v.image, pixels * 16, hipMemcpyDeviceToHost, ws->stream));
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
    auto to8 = [](float x) {
      x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
      return uint8_t(x * 255.0f + 0.5f);
    };
    for (size_t q = 0; q < pixels; ++q) {
      const float T = im[q * 4 + 3];
      for (int c = 0; c < 3; ++c) rgba8_out[q * 4 + c] = to8(im[q * 4 + c] + T * f->background[c]);
      rgba8_out[q * 4 + 3] = to8(1.0f - T);
    }
  }
  return 0;
}

}  // extern "C"
