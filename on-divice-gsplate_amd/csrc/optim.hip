// Per-slot update: the projection adjoint (2D splat gradient -> position,
// log scale, quaternion, opacity logit, SH DC and rest), the densification
// statistics, and Adam on all of them, fused in one pass
// (engine/shaders/train_step.glsl:62-240, train_step_sh.comp:21-44).  One
// thread per slot; the SH planes are coefficient-major, so the SH loop reads
// and writes each of sh, m, v as 768 B contiguous per wave and coefficient.
// HBM-bound: about 6 x (64 B + 12 B per SH coefficient) + 36 B per slot.
//
// Densification (train_densify_mark.comp, train_densify_spawn.comp) by stream
// compaction: the mark kernel flags prune / clone / split per slot, two
// exclusive scans rank the released slots (the free list) and the spawning
// ones, and the spawn kernel puts the k-th spawner into the k-th free slot or
// past the slots in use.  The slot each new gaussian lands in is therefore a
// function of the flags alone, not of atomic order.
#include "gs_common.h"
#include "gs_internal.h"

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
  const float* __restrict__ grad2d;
  const float4* __restrict__ records;
  const uint32_t* __restrict__ tiles;
  uint32_t capacity;
  uint32_t sh_coeffs;
};

__device__ __forceinline__ float adam1(float p, float g, float& m, float& v, float lr, const gs_hip_frame& f) {
  m = g * (1.0f - f.beta1) + m * f.beta1;
  v = g * g * (1.0f - f.beta2) + v * f.beta2;
  return p - lr * (m * f.bias1) / (sqrtf(v * f.bias2) + f.epsilon);
}

__device__ __forceinline__ float4 adam4(float4 p, float4 g, float4* m_ptr, float4* v_ptr, float4 lr,
                                        const gs_hip_frame& f) {
  float4 m = *m_ptr, v = *v_ptr;
  p.x = adam1(p.x, g.x, m.x, v.x, lr.x, f);
  p.y = adam1(p.y, g.y, m.y, v.y, lr.y, f);
  p.z = adam1(p.z, g.z, m.z, v.z, lr.z, f);
  p.w = adam1(p.w, g.w, m.w, v.w, lr.w, f);
  *m_ptr = m;
  *v_ptr = v;
  return p;
}

__global__ void __launch_bounds__(256) k_adam(gs_hip_frame f, AdamArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.slots) return;
  float4 p0 = a.params[i * 4u], p1 = a.params[i * 4u + 1u], q = a.params[i * 4u + 2u], p3 = a.params[i * 4u + 3u];
  if (p1.w < 0.5f) return;  // free slot

  float4 dp0 = make_float4(0, 0, 0, 0), dp1 = dp0, dq = dp0, dp3 = dp0;
  float dcol[3] = {0.0f, 0.0f, 0.0f};
  float dir[3] = {0.0f, 0.0f, 0.0f};
  bool drew = false;
  const uint32_t degree = f.sh_degree;
  const int coeffs = min(sh_rest_coeffs(degree), int(a.sh_coeffs));

  if (a.tiles[i] > 0u) {
    float g[9];
    for (int c = 0; c < 9; ++c) {
      g[c] = a.grad2d[size_t(c) * a.capacity + i];
      drew = drew || g[c] != 0.0f;
    }
    const float W = float(f.cam.width), H = float(f.cam.height);
    // Densification statistics: view-space positional gradient in NDC units.
    {
      float4 st = a.stats[i];
      const float gx = g[0] * (0.5f * W), gy = g[1] * (0.5f * H);
      st.x += sqrtf(gx * gx + gy * gy);
      st.y += 1.0f;
      st.z = fmaxf(st.z, float(__float_as_uint(a.records[size_t(i) * 3u + 2u].w) >> 3));
      a.stats[i] = st;
    }
    const float* V = f.cam.rotation;
    const float tx = V[0] * p0.x + V[1] * p0.y + V[2] * p0.z + f.cam.translation[0];
    const float ty = V[3] * p0.x + V[4] * p0.y + V[5] * p0.z + f.cam.translation[1];
    const float tz = V[6] * p0.x + V[7] * p0.y + V[8] * p0.z + f.cam.translation[2];
    const float fx = f.cam.fx, fy = f.cam.fy, cx = f.cam.cx, cy = f.cam.cy;
    const float inv_z = 1.0f / tz;
    const float lim_x = 1.3f * fmaxf(cx, W - cx) / fx, lim_y = 1.3f * fmaxf(cy, H - cy) / fy;
    const float rx = tx * inv_z, ry = ty * inv_z;
    const bool clamp_x = rx < -lim_x || rx > lim_x, clamp_y = ry < -lim_y || ry > lim_y;
    const float ux = fminf(fmaxf(rx, -lim_x), lim_x), uy = fminf(fmaxf(ry, -lim_y), lim_y);
    const float j00 = fx * inv_z, j02 = -fx * ux * inv_z, j11 = fy * inv_z, j12 = -fy * uy * inv_z;
    // T = J V (2x3).
    float T[2][3];
    for (int c = 0; c < 3; ++c) {
      T[0][c] = j00 * V[c] + j02 * V[6 + c];
      T[1][c] = j11 * V[3 + c] + j12 * V[6 + c];
    }
    const float qlen = fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
    const float qw = q.x / qlen, qx = q.y / qlen, qy = q.z / qlen, qz = q.w / qlen;
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
    // U = T Sigma (2x3); cov = U T^T.
    float U[2][3];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) U[r][c] = T[r][0] * Sig[0][c] + T[r][1] * Sig[1][c] + T[r][2] * Sig[2][c];
    const float A = U[0][0] * T[0][0] + U[0][1] * T[0][1] + U[0][2] * T[0][2] + kDilation;
    const float B = U[0][0] * T[1][0] + U[0][1] * T[1][1] + U[0][2] * T[1][2];
    const float C = U[1][0] * T[1][0] + U[1][1] * T[1][1] + U[1][2] * T[1][2] + kDilation;
    const float D = A * C - B * B;

    // Colour: clamped channels get no gradient; DC and the view direction.
    const float drx = p0.x - f.cam.position[0], dry = p0.y - f.cam.position[1], drz = p0.z - f.cam.position[2];
    const float dlen = fmaxf(sqrtf(drx * drx + dry * dry + drz * drz), 1e-12f);
    dir[0] = drx / dlen, dir[1] = dry / dlen, dir[2] = drz / dlen;
    const uint32_t clamped = __float_as_uint(a.records[size_t(i) * 3u + 2u].w) & 7u;
    dcol[0] = (clamped & 1u) ? 0.0f : g[6];
    dcol[1] = (clamped & 2u) ? 0.0f : g[7];
    dcol[2] = (clamped & 4u) ? 0.0f : g[8];
    const float b0 = 0.28209479177387814f;
    dp3.x = b0 * dcol[0], dp3.y = b0 * dcol[1], dp3.z = b0 * dcol[2];
    float ddir[3] = {0.0f, 0.0f, 0.0f};
    if (coeffs > 0) {
      float gx[16], gy[16], gz[16];
      sh_basis_grad(dir[0], dir[1], dir[2], degree, gx, gy, gz);
      for (int k = 0; k < coeffs; ++k) {
        const float* sk = a.sh + (size_t(k) * a.capacity + i) * 3u;
        const float w = sk[0] * dcol[0] + sk[1] * dcol[1] + sk[2] * dcol[2];
        ddir[0] += w * gx[k + 1];
        ddir[1] += w * gy[k + 1];
        ddir[2] += w * gz[k + 1];
      }
    }
    const float dd = dir[0] * ddir[0] + dir[1] * ddir[1] + dir[2] * ddir[2];
    const float dpos_dir[3] = {(ddir[0] - dir[0] * dd) / dlen, (ddir[1] - dir[1] * dd) / dlen,
                               (ddir[2] - dir[2] * dd) / dlen};

    // Opacity logit.
    const float op = 1.0f / (1.0f + __expf(-p0.w));
    dp0.w = g[5] * op * (1.0f - op);

    // Conic (C/D, -B/D, A/D) -> dilated covariance.
    const float ga = g[2], gb = g[3], gc = g[4];
    const float inv_d2 = 1.0f / (D * D);
    const float gA = (-C * C * ga + B * C * gb - B * B * gc) * inv_d2;
    const float gB = (2.0f * B * C * ga - (D + 2.0f * B * B) * gb + 2.0f * A * B * gc) * inv_d2;
    const float gC = (-B * B * ga + A * B * gb - A * A * gc) * inv_d2;
    const float G2[2][2] = {{gA, 0.5f * gB}, {0.5f * gB, gC}};
    // dSigma = T^T G2 T; dT = 2 G2 T Sigma.
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
    // dJ = dT V^T: only J00, J02, J11, J12 depend on t.
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
    // Screen mean.
    dtx += g[0] * fx * inv_z;
    dty += g[1] * fy * inv_z;
    dtz += -(g[0] * fx * tx + g[1] * fy * ty) * inv_z2;
    dp0.x = V[0] * dtx + V[3] * dty + V[6] * dtz + dpos_dir[0];
    dp0.y = V[1] * dtx + V[4] * dty + V[7] * dtz + dpos_dir[1];
    dp0.z = V[2] * dtx + V[5] * dty + V[8] * dtz + dpos_dir[2];

    // Sigma = M M^T -> dM = 2 dSigma M; M = R diag(s).
    float dM[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        dM[r][c] = 2.0f * (dSig[r][0] * M[0][c] + dSig[r][1] * M[1][c] + dSig[r][2] * M[2][c]);
    dp1.x = (dM[0][0] * R[0][0] + dM[1][0] * R[1][0] + dM[2][0] * R[2][0]) * s[0];
    dp1.y = (dM[0][1] * R[0][1] + dM[1][1] * R[1][1] + dM[2][1] * R[2][1]) * s[1];
    dp1.z = (dM[0][2] * R[0][2] + dM[1][2] * R[1][2] + dM[2][2] * R[2][2]) * s[2];
    float gR[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) gR[r][c] = dM[r][c] * s[c];
    const float dw = 2.0f * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] +
                             qx * gR[2][1]);
    const float dx = 2.0f * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2.0f * qx * gR[1][1] -
                             qw * gR[1][2] + qz * gR[2][0] + qw * gR[2][1] - 2.0f * qx * gR[2][2]);
    const float dy = 2.0f * (-2.0f * qy * gR[0][0] + qx * gR[0][1] + qw * gR[0][2] + qx * gR[1][0] +
                             qz * gR[1][2] - qw * gR[2][0] + qz * gR[2][1] - 2.0f * qy * gR[2][2]);
    const float dz = 2.0f * (-2.0f * qz * gR[0][0] - qw * gR[0][1] + qx * gR[0][2] + qw * gR[1][0] -
                             2.0f * qz * gR[1][1] + qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
    const float qd = qw * dw + qx * dx + qy * dy + qz * dz;
    dq.x = (dw - qw * qd) / qlen;
    dq.y = (dx - qx * qd) / qlen;
    dq.z = (dy - qy * qd) / qlen;
    dq.w = (dz - qz * qd) / qlen;
  }

  if (!drew && (f.flags & GS_HIP_FLAG_SPARSE_ADAM) != 0u) return;
  const float4 lr0 = make_float4(f.lr_position, f.lr_position, f.lr_position, f.lr_opacity);
  const float4 lr1 = make_float4(f.lr_scale, f.lr_scale, f.lr_scale, 0.0f);
  const float4 lrq = make_float4(f.lr_rotation, f.lr_rotation, f.lr_rotation, f.lr_rotation);
  const float4 lr3 = make_float4(f.lr_sh_dc, f.lr_sh_dc, f.lr_sh_dc, 0.0f);
  a.params[i * 4u] = adam4(p0, dp0, a.m_params + i * 4u, a.v_params + i * 4u, lr0, f);
  a.params[i * 4u + 1u] = adam4(p1, dp1, a.m_params + i * 4u + 1u, a.v_params + i * 4u + 1u, lr1, f);
  a.params[i * 4u + 2u] = adam4(q, dq, a.m_params + i * 4u + 2u, a.v_params + i * 4u + 2u, lrq, f);
  a.params[i * 4u + 3u] = adam4(p3, dp3, a.m_params + i * 4u + 3u, a.v_params + i * 4u + 3u, lr3, f);

  if (coeffs > 0) {
    float basis[16];
    sh_basis(dir[0], dir[1], dir[2], degree, basis);
    for (int k = 0; k < coeffs; ++k) {
      const size_t o = (size_t(k) * a.capacity + i) * 3u;
      for (int c = 0; c < 3; ++c) {
        float m = a.m_sh[o + c], v = a.v_sh[o + c];
        a.sh[o + c] = adam1(a.sh[o + c], basis[k + 1] * dcol[c], m, v, f.lr_sh_rest, f);
        a.m_sh[o + c] = m;
        a.v_sh[o + c] = v;
      }
    }
  }
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

extern "C" {

int gs_hip_adam(gs_hip_workspace* ws, const gs_hip_frame* f) {
  if (ws == nullptr || f == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "adam: null argument");
  if (f->slots == 0) return 0;
  AdamArgs a;
  a.params = reinterpret_cast<float4*>(ws->params);
  a.m_params = reinterpret_cast<float4*>(ws->m_params);
  a.v_params = reinterpret_cast<float4*>(ws->v_params);
  a.sh = ws->sh;
  a.m_sh = ws->m_sh;
  a.v_sh = ws->v_sh;
  a.stats = reinterpret_cast<float4*>(ws->stats);
  a.grad2d = ws->grad2d;
  a.records = reinterpret_cast<const float4*>(ws->records);
  a.tiles = ws->tiles;
  a.capacity = ws->capacity;
  a.sh_coeffs = ws->sh_coeffs;
  gsh::timer_begin(ws, "adam");
  k_adam<<<gsh::div_up(f->slots, 256), 256, 0, ws->stream>>>(*f, a);
  GS_CHECK_LAUNCH("k_adam");
  gsh::timer_end(ws);
  return 0;
}

int gs_hip_densify(gs_hip_workspace* ws, const gs_hip_frame* f, uint32_t* slots, uint32_t counts[3]) {
  if (ws == nullptr || f == nullptr || slots == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "densify: null");
  if (f->slots != *slots) return gsh::set_error(GS_HIP_E_ARGUMENT, "densify: frame slots differ from *slots");
  if (counts != nullptr) counts[0] = counts[1] = counts[2] = 0;
  const uint32_t n = *slots;
  if (n == 0) return 0;
  DensifyArgs a;
  a.params = reinterpret_cast<float4*>(ws->params);
  a.m_params = reinterpret_cast<float4*>(ws->m_params);
  a.v_params = reinterpret_cast<float4*>(ws->v_params);
  a.sh = ws->sh;
  a.m_sh = ws->m_sh;
  a.v_sh = ws->v_sh;
  a.stats = reinterpret_cast<float4*>(ws->stats);
  // Per-step scratch reused: the next step's projection rewrites all of these.
  a.action = ws->flags;
  a.release = ws->depth_keys;
  a.spawn = ws->depth_order;
  a.free_rank = ws->offsets;
  a.spawn_rank = ws->scan_tmp;
  a.free_list = ws->tiles;
  a.totals = ws->totals;
  a.capacity = ws->capacity;
  a.sh_coeffs = ws->sh_coeffs;
  const uint32_t blocks = gsh::div_up(n, 256);
  gsh::timer_begin(ws, "densify");
  k_zero_counters<<<1, 64, 0, ws->stream>>>(ws->totals);
  GS_CHECK_LAUNCH("k_zero_counters");
  k_densify_mark<<<blocks, 256, 0, ws->stream>>>(*f, a);
  GS_CHECK_LAUNCH("k_densify_mark");
  int rc = gsh::exclusive_scan(ws, a.release, nullptr, a.free_rank, n, ws->totals + 6);
  if (rc == 0) rc = gsh::exclusive_scan(ws, a.spawn, nullptr, a.spawn_rank, n, ws->totals + 7);
  if (rc != 0) return rc;
  k_free_list<<<blocks, 256, 0, ws->stream>>>(n, a);
  GS_CHECK_LAUNCH("k_free_list");
  k_densify_spawn<<<blocks, 256, 0, ws->stream>>>(*f, a);
  GS_CHECK_LAUNCH("k_densify_spawn");
  k_zero_stats<<<gsh::div_up(ws->capacity, 256), 256, 0, ws->stream>>>(a.stats, ws->capacity);
  GS_CHECK_LAUNCH("k_zero_stats");
  GS_HIP_TRY(hipMemcpyAsync(ws->host_totals, ws->totals, 32, hipMemcpyDeviceToHost, ws->stream));
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
  int rc = gs_hip_project(ws, f);
  if (rc == 0) rc = gs_hip_bin(ws, f, pairs);
  if (rc == 0) rc = gs_hip_forward(ws, f);
  if (rc == 0) rc = gs_hip_loss(ws, f, losses);
  if (rc == 0) rc = gs_hip_backward(ws, f);
  if (rc == 0) rc = gs_hip_adam(ws, f);
  return rc;
}

int gs_hip_render(gs_hip_workspace* ws, const gs_hip_frame* f, uint8_t* rgba8_out) {
  int rc = gs_hip_project(ws, f);
  if (rc == 0) rc = gs_hip_bin(ws, f, nullptr);
  if (rc == 0) rc = gs_hip_forward(ws, f);
  if (rc != 0) return rc;
  if (rgba8_out != nullptr) {
    const size_t pixels = size_t(ws->width) * ws->height;
    std::vector<float> im(pixels * 4);
    GS_HIP_TRY(hipMemcpyAsync(im.data(), ws->image, pixels * 16, hipMemcpyDeviceToHost, ws->stream));
    GS_HIP_TRY(hipStreamSynchronize(ws->stream));
    auto to8 = [](float v) {
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      return uint8_t(v * 255.0f + 0.5f);
    };
    for (size_t p = 0; p < pixels; ++p) {
      const float T = im[p * 4 + 3];
      for (int c = 0; c < 3; ++c) rgba8_out[p * 4 + c] = to8(im[p * 4 + c] + T * f->background[c]);
      rgba8_out[p * 4 + 3] = to8(1.0f - T);
    }
  }
  return 0;
}

}  // extern "C"
