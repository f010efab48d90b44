// Scene files on the device: SPZ v2/v3 decoding into a workspace and 3DGS PLY
// rows of a workspace's alive slots, the on-disk formats either side of the
// training step at scale (a million splats is 250 MB of PLY rows).
//
// SPZ (engine/src/splat_scene.cpp:367-448 decodes it on the host): the gzip
// stream is inflated on the host (zlib); the fixed-point attribute arrays are
// unpacked by one thread per splat straight into the slot layout and the SH
// planes.  Every value equals the host decoder's (gs::load_spz) bit for bit:
// the same IEEE operations in the same order with contraction off, and the
// opacity logit of each byte taken from a 256-entry table the host computes
// with its own log.
//
// PLY export (trainer.cpp:1773-1800 export_scene, splat_scene.cpp:336-365
// save_ply, convert_convention): the alive slots are ranked by an exclusive
// scan; one thread per alive slot writes its row (x y z, nx ny nz = 0, f_dc,
// f_rest channel-major, opacity, scale, rot) with the quaternion normalised
// and, for the RDF convention, the half turn about x applied (y and z
// negated, the quaternion premultiplied by (0, 1, 0, 0), each SH coefficient's
// sign by its y/z parity).  The rows equal gs::save_ply's after the same
// conversions on the host.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs_internal.h"

namespace {

constexpr uint32_t kSpzMagic = 0x5053474eu;  // "NGSP"
constexpr float kSpzColourScale = 0.15f;

struct SpzLayout {
  uint32_t n, version, file_coeffs, coeffs, frac_bits;
  uint64_t o_pos, o_alpha, o_col, o_scale, o_rot, o_sh;
};

__global__ void __launch_bounds__(256) k_spz_unpack(const uint8_t* __restrict__ raw, const float* __restrict__ logit,
                                                    SpzLayout l, float4* __restrict__ params, float* __restrict__ sh,
                                                    uint32_t capacity, uint32_t ws_coeffs) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= l.n) return;
  const float pos_scale = 1.0f / float(1u << l.frac_bits);
  float pos[3], dc[3], ls[3];
  for (int k = 0; k < 3; ++k) {
    const uint8_t* p = raw + l.o_pos + (size_t(i) * 3 + k) * 3;
    int32_t fixed = int32_t(p[0]) | (int32_t(p[1]) << 8) | (int32_t(p[2]) << 16);
    if (fixed & 0x800000) fixed |= int32_t(0xff000000u);  // sign-extend 24 bits
    pos[k] = float(fixed) * pos_scale;
    dc[k] = (float(raw[l.o_col + size_t(i) * 3 + k]) / 255.0f - 0.5f) / kSpzColourScale;
    ls[k] = float(raw[l.o_scale + size_t(i) * 3 + k]) / 16.0f - 10.0f;
  }
  const float opacity = logit[raw[l.o_alpha + i]];
  float xyzw[4];
  if (l.version >= 3) {
    // "smallest three": 2 bits name the largest component; the other three
    // are 9-bit magnitudes (scaled by 1/sqrt 2) with a sign bit, packed from
    // the last component down; the largest comes from unit norm.
    const uint8_t* r = raw + l.o_rot + size_t(i) * 4;
    uint32_t bits = uint32_t(r[0]) | (uint32_t(r[1]) << 8) | (uint32_t(r[2]) << 16) | (uint32_t(r[3]) << 24);
    const int largest = int(bits >> 30);
    const uint32_t mask = (1u << 9) - 1u;
    float sum = 0.0f;
    for (int c = 3; c >= 0; --c) {
      if (c == largest) continue;
      const float v = 0.70710678f * float(bits & mask) / float(mask);
      xyzw[c] = ((bits >> 9) & 1u) ? -v : v;
      bits >>= 10;
      sum += xyzw[c] * xyzw[c];
    }
    xyzw[largest] = sqrtf(fmaxf(0.0f, 1.0f - sum));
  } else {
    const uint8_t* r = raw + l.o_rot + size_t(i) * 3;
    for (int k = 0; k < 3; ++k) xyzw[k] = float(r[k]) / 127.5f - 1.0f;
    xyzw[3] = sqrtf(fmaxf(0.0f, 1.0f - (xyzw[0] * xyzw[0] + xyzw[1] * xyzw[1] + xyzw[2] * xyzw[2])));
  }
  params[size_t(i) * 4u] = make_float4(pos[0], pos[1], pos[2], opacity);
  params[size_t(i) * 4u + 1u] = make_float4(ls[0], ls[1], ls[2], 1.0f);
  params[size_t(i) * 4u + 2u] = make_float4(xyzw[3], xyzw[0], xyzw[1], xyzw[2]);
  params[size_t(i) * 4u + 3u] = make_float4(dc[0], dc[1], dc[2], 0.0f);
  for (uint32_t j = 0; j < ws_coeffs; ++j) {
    float* o = sh + (size_t(j) * capacity + i) * 3u;
    for (int c = 0; c < 3; ++c) {
      const float v = j < l.coeffs ? (float(raw[l.o_sh + (size_t(i) * l.file_coeffs + j) * 3 + c]) - 128.0f) / 128.0f
                                   : 0.0f;
      o[c] = v;
    }
  }
}

__global__ void __launch_bounds__(256) k_alive(const float4* __restrict__ params, uint32_t slots,
                                               uint32_t* __restrict__ alive) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < slots) alive[i] = params[size_t(i) * 4u + 1u].w >= 0.5f ? 1u : 0u;
}

// Sign of SH rest coefficient j under the half turn about x: the parity of
// its basis function's y and z factors.
__constant__ float kHalfTurnSign[GS_HIP_MAX_SH_REST] = {-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};

__global__ void __launch_bounds__(256) k_export_rows(const float4* __restrict__ params, const float* __restrict__ sh,
                                                     const uint32_t* __restrict__ alive,
                                                     const uint32_t* __restrict__ rank, uint32_t slots,
                                                     uint32_t capacity, uint32_t coeffs, uint32_t rdf,
                                                     float* __restrict__ rows) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots || alive[i] == 0u) return;
  const uint32_t per = 17u + 3u * coeffs;
  float* r = rows + size_t(rank[i]) * per;
  const float4 p0 = params[size_t(i) * 4u], p1 = params[size_t(i) * 4u + 1u];
  const float4 q = params[size_t(i) * 4u + 2u], p3 = params[size_t(i) * 4u + 3u];
  const float len = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  float w = 1.0f, x = 0.0f, y = 0.0f, z = 0.0f;
  if (len > 0.0f) w = q.x / len, x = q.y / len, y = q.z / len, z = q.w / len;
  r[0] = p0.x;
  r[1] = rdf ? -p0.y : p0.y;
  r[2] = rdf ? -p0.z : p0.z;
  r[3] = r[4] = r[5] = 0.0f;
  r[6] = p3.x;
  r[7] = p3.y;
  r[8] = p3.z;
  for (int c = 0; c < 3; ++c) {
    for (uint32_t j = 0; j < coeffs; ++j) {
      const float v = sh[(size_t(j) * capacity + i) * 3u + c];
      r[9 + c * coeffs + j] = rdf ? v * kHalfTurnSign[j] : v;
    }
  }
  float* t = r + 9 + 3 * coeffs;
  t[0] = p0.w;
  t[1] = p1.x;
  t[2] = p1.y;
  t[3] = p1.z;
  if (rdf) {
    t[4] = -x, t[5] = w, t[6] = -z, t[7] = y;
  } else {
    t[4] = w, t[5] = x, t[6] = y, t[7] = z;
  }
}

// Inflates the gzip stream; with header_only, stops after the first 16 bytes.
int gunzip(const uint8_t* data, size_t size, bool header_only, std::vector<uint8_t>& out) {
  if (size < 2 || data[0] != 0x1f || data[1] != 0x8b) return gsh::set_error(GS_HIP_E_IO, "SPZ: not a gzip stream");
  z_stream zs{};
  zs.next_in = const_cast<Bytef*>(data);
  zs.avail_in = static_cast<uInt>(std::min<size_t>(size, 0xffffffffu));
  if (inflateInit2(&zs, 16 + MAX_WBITS) != Z_OK) return gsh::set_error(GS_HIP_E_IO, "SPZ: gzip init failed");
  out.clear();
  std::vector<uint8_t> chunk(header_only ? 16 : 1 << 20);
  int rc = Z_OK;
  while (rc == Z_OK) {
    zs.next_out = chunk.data();
    zs.avail_out = static_cast<uInt>(chunk.size());
    rc = inflate(&zs, Z_NO_FLUSH);
    if (rc != Z_OK && rc != Z_STREAM_END) break;
    out.insert(out.end(), chunk.data(), chunk.data() + (chunk.size() - zs.avail_out));
    if (header_only && out.size() >= 16) break;
  }
  inflateEnd(&zs);
  if (header_only ? out.size() < 16 : rc != Z_STREAM_END) {
    return gsh::set_error(GS_HIP_E_IO, "SPZ: decompression failed (corrupt file?)");
  }
  return 0;
}

template <typename T>
T load_le(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

int degree_of(uint32_t degree_byte) { return degree_byte > 4u ? -1 : int(degree_byte); }

int parse_header(const std::vector<uint8_t>& raw, uint32_t* count, int* degree, int* antialiased, uint32_t* version,
                 uint32_t* frac_bits) {
  if (raw.size() < 16 || load_le<uint32_t>(raw.data()) != kSpzMagic) return gsh::set_error(GS_HIP_E_IO, "SPZ: bad header");
  *version = load_le<uint32_t>(raw.data() + 4);
  if (*version < 2 || *version > 3) return gsh::set_error(GS_HIP_E_IO, "SPZ: only versions 2 and 3 are supported");
  *count = load_le<uint32_t>(raw.data() + 8);
  *degree = degree_of(raw[12]);
  if (*degree < 0) return gsh::set_error(GS_HIP_E_IO, "SPZ: bad SH degree");
  *frac_bits = raw[13];
  if (*frac_bits > 24u) return gsh::set_error(GS_HIP_E_IO, "SPZ: bad fixed-point fraction bits");
  *antialiased = (raw[14] & 1u) != 0 ? 1 : 0;
  if (*count == 0) return gsh::set_error(GS_HIP_E_IO, "SPZ: no splats");
  return 0;
}

int rest_coeffs(int degree) { return degree <= 0 ? 0 : (degree + 1) * (degree + 1) - 1; }

}  // namespace

extern "C" {

int gs_hip_spz_info(const uint8_t* spz, size_t bytes, uint32_t* count, int* sh_degree, int* antialiased) {
  gsh::clear_error();
  if (spz == nullptr || count == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "spz_info: null argument");
  if (bytes >= 4 && load_le<uint32_t>(spz) == kSpzMagic) {
    return gsh::set_error(GS_HIP_E_IO, "SPZ v4 (zstd) is not supported (v2/v3 only)");
  }
  std::vector<uint8_t> raw;
  int rc = gunzip(spz, bytes, true, raw);
  uint32_t version = 0, frac = 0;
  int degree = 0, aa = 0;
  if (rc == 0) rc = parse_header(raw, count, &degree, &aa, &version, &frac);
  if (rc != 0) return rc;
  if (sh_degree != nullptr) *sh_degree = std::min(degree, 3);
  if (antialiased != nullptr) *antialiased = aa;
  return 0;
}

int gs_hip_workspace_load_spz(gs_hip_workspace* ws, const uint8_t* spz, size_t bytes, uint32_t* count,
                              int* sh_degree, int* antialiased) {
  gsh::clear_error();
  if (ws == nullptr || spz == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "load_spz: null argument");
  if (bytes >= 4 && load_le<uint32_t>(spz) == kSpzMagic) {
    return gsh::set_error(GS_HIP_E_IO, "SPZ v4 (zstd) is not supported (v2/v3 only)");
  }
  std::vector<uint8_t> raw;
  int rc = gunzip(spz, bytes, false, raw);
  uint32_t n = 0, version = 0, frac = 0;
  int degree = 0, aa = 0;
  if (rc == 0) rc = parse_header(raw, &n, &degree, &aa, &version, &frac);
  if (rc != 0) return rc;
  if (n > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "load_spz: more splats than the workspace holds");
  SpzLayout l{};
  l.n = n;
  l.version = version;
  l.file_coeffs = uint32_t(rest_coeffs(degree));
  l.coeffs = std::min<uint32_t>(uint32_t(rest_coeffs(std::min(degree, 3))), ws->sh_coeffs);
  l.frac_bits = frac;
  const uint64_t rot_bytes = version >= 3 ? 4 : 3;
  l.o_pos = 16;
  l.o_alpha = l.o_pos + uint64_t(n) * 9;
  l.o_col = l.o_alpha + n;
  l.o_scale = l.o_col + uint64_t(n) * 3;
  l.o_rot = l.o_scale + uint64_t(n) * 3;
  l.o_sh = l.o_rot + uint64_t(n) * rot_bytes;
  if (raw.size() < l.o_sh + uint64_t(n) * l.file_coeffs * 3) return gsh::set_error(GS_HIP_E_IO, "SPZ: data truncated");
  // The opacity logit of each byte, with the host's log (gs::load_spz: p clamped to [0.5, 254.5] / 255).
  float logit[256];
  for (int b = 0; b < 256; ++b) {
    const float p = std::min(std::max(float(b) / 255.0f, 0.5f / 255.0f), 1.0f - 0.5f / 255.0f);
    logit[b] = std::log(p / (1.0f - p));
  }
  const size_t table = sizeof(logit), payload = raw.size();
  void* dev = nullptr;
  GS_HIP_TRY(hipMalloc(&dev, table + payload));
  float* d_logit = static_cast<float*>(dev);
  uint8_t* d_raw = static_cast<uint8_t*>(dev) + table;
  // Zero the model first, as gs_hip_workspace_load does (moments, statistics, unused slots).
  const size_t cap = ws->capacity, shn = cap * ws->sh_coeffs * 3;
  for (void* p : {(void*)ws->params, (void*)ws->m_params, (void*)ws->v_params}) {
    if (rc == 0) rc = gsh::fill_u32(ws, p, cap * 16, 0);
  }
  for (void* p : {(void*)ws->sh, (void*)ws->m_sh, (void*)ws->v_sh}) {
    if (rc == 0) rc = gsh::fill_u32(ws, p, shn, 0);
  }
  if (rc == 0) rc = gsh::fill_u32(ws, ws->stats, cap * 4, 0);
  hipError_t e = rc == 0 ? hipMemcpyAsync(d_logit, logit, table, hipMemcpyHostToDevice, ws->stream) : hipSuccess;
  if (e == hipSuccess && rc == 0) e = hipMemcpyAsync(d_raw, raw.data(), payload, hipMemcpyHostToDevice, ws->stream);
  if (e == hipSuccess && rc == 0) {
    k_spz_unpack<<<gsh::div_up(n, 256), 256, 0, ws->stream>>>(d_raw, d_logit, l, reinterpret_cast<float4*>(ws->params),
                                                              ws->sh, ws->capacity, ws->sh_coeffs);
    e = hipGetLastError();
  }
  const hipError_t s = hipStreamSynchronize(ws->stream);
  hipFree(dev);
  if (rc != 0) return rc;
  if (e != hipSuccess) return gsh::hip_error(e, "load_spz");
  if (s != hipSuccess) return gsh::hip_error(s, "load_spz sync");
  if (count != nullptr) *count = n;
  if (sh_degree != nullptr) *sh_degree = std::min(degree, 3);
  if (antialiased != nullptr) *antialiased = aa;
  return 0;
}

int gs_hip_export_ply(gs_hip_workspace* ws, uint32_t slots, int rdf, uint8_t* rows, size_t bytes, uint32_t* count) {
  gsh::clear_error();
  if (ws == nullptr || count == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "export_ply: null argument");
  if (slots > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "export_ply: more slots than capacity");
  gsh::View& v = gsh::view0(ws);
  const uint32_t per = 17u + 3u * ws->sh_coeffs;
  *count = 0;
  if (slots == 0) return 0;
  k_alive<<<gsh::div_up(slots, 256), 256, 0, ws->stream>>>(reinterpret_cast<const float4*>(ws->params), slots, v.flags);
  GS_CHECK_LAUNCH("k_alive");
  int rc = gsh::exclusive_scan(ws, v, v.flags, nullptr, v.scan_tmp, slots, v.scan_tmp + slots);
  if (rc != 0) return rc;
  uint32_t alive = 0;
  GS_HIP_TRY(hipMemcpyAsync(&alive, v.scan_tmp + slots, 4, hipMemcpyDeviceToHost, ws->stream));
  GS_HIP_TRY(hipStreamSynchronize(ws->stream));
  *count = alive;
  if (rows == nullptr) return 0;  // size query
  const size_t need = size_t(alive) * per * 4;
  if (bytes < need) return gsh::set_error(GS_HIP_E_CAPACITY, "export_ply: the row buffer is too small");
  if (alive == 0) return 0;
  void* dev = nullptr;
  GS_HIP_TRY(hipMalloc(&dev, need));
  k_export_rows<<<gsh::div_up(slots, 256), 256, 0, ws->stream>>>(
      reinterpret_cast<const float4*>(ws->params), ws->sh, v.flags, v.scan_tmp, slots, ws->capacity, ws->sh_coeffs,
      rdf != 0 ? 1u : 0u, static_cast<float*>(dev));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(rows, dev, need, hipMemcpyDeviceToHost, ws->stream);
  const hipError_t s = hipStreamSynchronize(ws->stream);
  hipFree(dev);
  if (e != hipSuccess) return gsh::hip_error(e, "export_ply");
  if (s != hipSuccess) return gsh::hip_error(s, "export_ply sync");
  return 0;
}

}  // extern "C"
