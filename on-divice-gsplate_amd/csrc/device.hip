// Device, memory and workspace entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "gs_common.h"
#include "gs_internal.h"

namespace gsh {

namespace {
thread_local int g_code = 0;
thread_local std::string g_message;
}  // namespace

int set_error(int code, const std::string& message) {
  g_code = code;
  g_message = message;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  return set_error(GS_HIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void clear_error() {
  g_code = 0;
  g_message.clear();
}

void timer_begin(gs_hip_workspace* ws, View& v, const char* name) {
  Timer& t = ws->timer;
  if (!t.enabled) return;
  if (t.used == t.start.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    t.start.push_back(a);
    t.stop.push_back(b);
    t.names.emplace_back();
  }
  t.names[t.used] = name;
  hipEventRecord(t.start[t.used], v.stream);
}

void timer_end(gs_hip_workspace* ws, View& v) {
  Timer& t = ws->timer;
  if (!t.enabled || t.used >= t.stop.size()) return;
  hipEventRecord(t.stop[t.used], v.stream);
  ++t.used;
}

}  // namespace gsh

namespace {

// Fills `count` u32 with `value` (a kernel, not hipMemsetAsync, so the step
// stays capturable in a graph).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t count, uint32_t value) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

template <typename T>
int alloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return 0;
}

template <typename T>
void release(T*& p) {
  if (p != nullptr) hipFree(p);
  p = nullptr;
}

int zero(gsh::View& v, void* p, uint64_t bytes) {
  if (bytes == 0) return 0;
  const uint64_t words = bytes / 4;
  if (words > 0) {
    k_fill_u32<<<gsh::div_up(words, 256), 256, 0, v.stream>>>(static_cast<uint32_t*>(p), words, 0u);
    GS_CHECK_LAUNCH("k_fill_u32");
  }
  return 0;
}

}  // namespace

namespace gsh {

int create_view(gs_hip_workspace* ws, View** out) {
  *out = nullptr;
  auto* v = new View();
  const uint64_t cap = std::max<uint32_t>(ws->capacity, 1);
  const uint64_t pixels = uint64_t(ws->width) * ws->height;
  const uint64_t tiles = uint64_t(ws->tiles_x) * ws->tiles_y;
  int rc = 0;
  auto A = [&](int r) {
    if (rc == 0) rc = r;
  };
  A(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) == hipSuccess
        ? 0
        : set_error(GS_HIP_E_HIP, "hipStreamCreate failed"));
  A(hipEventCreateWithFlags(&v->done, hipEventDisableTiming) == hipSuccess
        ? 0
        : set_error(GS_HIP_E_HIP, "hipEventCreate failed"));
  A(alloc(&v->records, cap * GS_HIP_RECORD_FLOATS));
  A(alloc(&v->footprint, cap * 4));
  A(alloc(&v->tiles, cap));
  A(alloc(&v->misc, cap));
  A(alloc(&v->depth_keys, cap));
  A(alloc(&v->depth_order, cap));
  A(alloc(&v->offsets, cap));
  A(alloc(&v->grad2d, cap * GS_HIP_GRAD2D_FLOATS));
  A(alloc(&v->sh_input, cap * 8));
  A(alloc(&v->flags, cap));
  A(alloc(&v->scan_tmp, cap + 1));
  A(alloc(&v->ranges, (tiles + 1) * 2));
  A(alloc(&v->tile_order, tiles));
  A(alloc(&v->image, pixels * 4));
  A(alloc(&v->last, pixels));
  A(alloc(&v->target, pixels));
  A(alloc(&v->dldc, pixels * 4));
  A(alloc(&v->ssim_maps, pixels * 9));
  A(alloc(&v->tile_loss, tiles * 4));
  A(alloc(&v->totals, 16));
  v->block_sums_entries = div_up(cap + 1, 4096) + 1024;
  A(alloc(&v->block_sums, v->block_sums_entries));
  if (rc == 0 && hipHostMalloc(reinterpret_cast<void**>(&v->host_totals), 64) != hipSuccess) {
    rc = set_error(GS_HIP_E_HIP, "hipHostMalloc failed");
  }
  if (rc == 0 && hipHostMalloc(reinterpret_cast<void**>(&v->host_tile_loss), tiles * 16) != hipSuccess) {
    rc = set_error(GS_HIP_E_HIP, "hipHostMalloc failed");
  }
  if (rc == 0) rc = reserve_sort(*v, cap);
  if (rc == 0) rc = zero(*v, v->totals, 64);
  if (rc == 0) rc = zero(*v, v->target, pixels * 4);
  if (rc == 0 && hipStreamSynchronize(v->stream) != hipSuccess) rc = set_error(GS_HIP_E_HIP, "view init sync");
  if (rc != 0) {
    destroy_view(v);
    return rc;
  }
  *out = v;
  return 0;
}

void destroy_view(View* v) {
  if (v == nullptr) return;
  if (v->stream != nullptr) hipStreamSynchronize(v->stream);
  release(v->records);
  release(v->footprint);
  release(v->tiles);
  release(v->misc);
  release(v->depth_keys);
  release(v->depth_order);
  release(v->offsets);
  release(v->grad2d);
  release(v->sh_input);
  release(v->flags);
  release(v->scan_tmp);
  release(v->pair_tiles);
  release(v->pair_gaussians);
  release(v->ranges);
  release(v->tile_order);
  release(v->image);
  release(v->last);
  release(v->target);
  release(v->dldc);
  release(v->ssim_maps);
  release(v->tile_loss);
  release(v->totals);
  release(v->block_sums);
  release(v->sort.keys_alt);
  release(v->sort.vals_alt);
  release(v->sort.hist);
  release(v->sort.digit_totals);
  if (v->host_totals != nullptr) hipHostFree(v->host_totals);
  if (v->host_tile_loss != nullptr) hipHostFree(v->host_tile_loss);
  if (v->done != nullptr) hipEventDestroy(v->done);
  if (v->stream != nullptr) hipStreamDestroy(v->stream);
  delete v;
}

int reserve_pairs(gs_hip_workspace* ws, View& v, uint64_t pairs) {
  (void)ws;
  if (pairs <= v.pair_capacity) return 0;
  if (pairs >= (uint64_t(1) << 32)) return set_error(GS_HIP_E_CAPACITY, "reserve_pairs: more than 2^32 pairs");
  // Headroom: views differ in pair count (trainer.cpp:588-603 keeps 1/8 + 1/4).
  uint64_t n = std::min<uint64_t>(pairs + pairs / 4, (uint64_t(1) << 32) - 1);
  n = std::max<uint64_t>(n, 1 << 16);
  GS_HIP_TRY(hipStreamSynchronize(v.stream));
  release(v.pair_tiles);
  release(v.pair_gaussians);
  v.pair_capacity = 0;
  int rc = alloc(&v.pair_tiles, n);
  if (rc == 0) rc = alloc(&v.pair_gaussians, n);
  if (rc == 0) rc = reserve_sort(v, n);
  if (rc != 0) return rc;
  v.pair_capacity = n;
  return 0;
}

}  // namespace gsh

extern "C" {

int gs_hip_init(int device) {
  gsh::clear_error();
  int count = 0;
  GS_HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return gsh::set_error(GS_HIP_E_ARGUMENT, "gs_hip_init: no such device");
  GS_HIP_TRY(hipSetDevice(device));
  return 0;
}

int gs_hip_device_name(char* buffer, size_t size) {
  int dev = 0;
  GS_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GS_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (buffer != nullptr && size > 0) {
    std::snprintf(buffer, size, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  return 0;
}

int gs_hip_malloc(void** ptr, size_t bytes) {
  if (ptr == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "gs_hip_malloc: null ptr");
  GS_HIP_TRY(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return 0;
}

int gs_hip_free(void* ptr) {
  if (ptr != nullptr) GS_HIP_TRY(hipFree(ptr));
  return 0;
}

// The plain copies wait for all device work first: workspaces run on their own
// non-blocking streams, which a null-stream copy would not order against.
int gs_hip_upload(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_download(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int gs_hip_memset(void* dst, int value, size_t bytes) {
  if (bytes == 0) return 0;
  GS_HIP_TRY(hipDeviceSynchronize());
  GS_HIP_TRY(hipMemset(dst, value, bytes));
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_synchronize(void) {
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_last_error(char* buffer, size_t size) {
  if (buffer != nullptr && size > 0) std::snprintf(buffer, size, "%s", gsh::g_message.c_str());
  return gsh::g_code;
}

int gs_hip_workspace_create(gs_hip_workspace** out, uint32_t capacity, int sh_degree, uint32_t width,
                            uint32_t height) {
  gsh::clear_error();
  if (out == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: null handle");
  *out = nullptr;
  if (sh_degree < 0 || sh_degree > 3) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: sh_degree 0..3");
  if (width == 0 || height == 0 || width > 32768 || height > 32768) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_create: image size out of range");
  }
  if (capacity > (1u << 30)) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_create: capacity too large");
  auto* ws = new gs_hip_workspace();
  ws->capacity = capacity;
  ws->sh_degree = sh_degree;
  ws->sh_coeffs = sh_degree == 0 ? 0u : uint32_t((sh_degree + 1) * (sh_degree + 1) - 1);
  ws->width = width;
  ws->height = height;
  ws->tiles_x = gsh::div_up(width, GS_HIP_TILE);
  ws->tiles_y = gsh::div_up(height, GS_HIP_TILE);
  const uint64_t cap = std::max<uint32_t>(capacity, 1);
  const uint64_t shn = cap * ws->sh_coeffs * 3;
  int rc = 0;
  auto A = [&](int r) {
    if (rc == 0) rc = r;
  };
  A(alloc(&ws->params, cap * 16));
  A(alloc(&ws->sh, shn));
  A(alloc(&ws->m_params, cap * 16));
  A(alloc(&ws->v_params, cap * 16));
  A(alloc(&ws->m_sh, shn));
  A(alloc(&ws->v_sh, shn));
  A(alloc(&ws->stats, cap * 4));
  gsh::View* v = nullptr;
  if (rc == 0) rc = gsh::create_view(ws, &v);
  if (v != nullptr) ws->views.push_back(v);
  if (rc == 0) {
    for (void* p : {(void*)ws->params, (void*)ws->m_params, (void*)ws->v_params}) A(zero(*v, p, cap * 64));
    for (void* p : {(void*)ws->sh, (void*)ws->m_sh, (void*)ws->v_sh}) A(zero(*v, p, shn * 4));
    A(zero(*v, ws->stats, cap * 16));
    if (rc == 0 && hipStreamSynchronize(v->stream) != hipSuccess) A(gsh::set_error(GS_HIP_E_HIP, "workspace init sync"));
  }
  if (rc != 0) {
    gs_hip_workspace_destroy(ws);
    return rc;
  }
  *out = ws;
  return 0;
}

int gs_hip_workspace_views(gs_hip_workspace* ws, uint32_t count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_views: null workspace");
  if (count == 0 || count > GS_HIP_MAX_VIEWS) {
    return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_views: 1 to " + std::to_string(GS_HIP_MAX_VIEWS) + " views");
  }
  while (ws->views.size() < count) {
    gsh::View* v = nullptr;
    const int rc = gsh::create_view(ws, &v);
    if (rc != 0) return rc;
    ws->views.push_back(v);
  }
  return 0;
}

int gs_hip_workspace_destroy(gs_hip_workspace* ws) {
  if (ws == nullptr) return 0;
  hipDeviceSynchronize();
  for (gsh::View* v : ws->views) gsh::destroy_view(v);
  ws->views.clear();
  release(ws->params);
  release(ws->sh);
  release(ws->m_params);
  release(ws->v_params);
  release(ws->m_sh);
  release(ws->v_sh);
  release(ws->stats);
  for (hipEvent_t e : ws->timer.start) hipEventDestroy(e);
  for (hipEvent_t e : ws->timer.stop) hipEventDestroy(e);
  delete ws;
  return 0;
}

int gs_hip_workspace_view_buffer(gs_hip_workspace* ws, uint32_t view, int which, void** ptr, size_t* bytes) {
  if (ws == nullptr || ptr == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: null");
  if (view >= ws->views.size()) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: no such view");
  const gsh::View& v = *ws->views[view];
  const size_t cap = ws->capacity, pixels = size_t(ws->width) * ws->height;
  const size_t tiles = size_t(ws->tiles_x) * ws->tiles_y, shn = cap * ws->sh_coeffs * 3;
  void* p = nullptr;
  size_t b = 0;
  switch (which) {
    case GS_HIP_BUF_PARAMS: p = ws->params, b = cap * 64; break;
    case GS_HIP_BUF_SH: p = ws->sh, b = shn * 4; break;
    case GS_HIP_BUF_M_PARAMS: p = ws->m_params, b = cap * 64; break;
    case GS_HIP_BUF_V_PARAMS: p = ws->v_params, b = cap * 64; break;
    case GS_HIP_BUF_M_SH: p = ws->m_sh, b = shn * 4; break;
    case GS_HIP_BUF_V_SH: p = ws->v_sh, b = shn * 4; break;
    case GS_HIP_BUF_STATS: p = ws->stats, b = cap * 16; break;
    case GS_HIP_BUF_RECORDS: p = v.records, b = cap * GS_HIP_RECORD_FLOATS * 4; break;
    case GS_HIP_BUF_FOOTPRINT: p = v.footprint, b = cap * 16; break;
    case GS_HIP_BUF_TILES: p = v.tiles, b = cap * 4; break;
    case GS_HIP_BUF_DEPTH_KEYS: p = v.depth_keys, b = cap * 4; break;
    case GS_HIP_BUF_DEPTH_ORDER: p = v.depth_order, b = cap * 4; break;
    case GS_HIP_BUF_OFFSETS: p = v.offsets, b = cap * 4; break;
    case GS_HIP_BUF_GRAD2D: p = v.grad2d, b = cap * GS_HIP_GRAD2D_FLOATS * 4; break;
    case GS_HIP_BUF_PAIR_TILES: p = v.pair_tiles, b = v.pair_capacity * 4; break;
    case GS_HIP_BUF_PAIR_GAUSSIANS: p = v.pair_gaussians, b = v.pair_capacity * 4; break;
    case GS_HIP_BUF_RANGES: p = v.ranges, b = (tiles + 1) * 8; break;
    case GS_HIP_BUF_IMAGE: p = v.image, b = pixels * 16; break;
    case GS_HIP_BUF_LAST: p = v.last, b = pixels * 4; break;
    case GS_HIP_BUF_TARGET: p = v.target, b = pixels * 4; break;
    case GS_HIP_BUF_DLDC: p = v.dldc, b = pixels * 16; break;
    case GS_HIP_BUF_TILE_LOSS: p = v.tile_loss, b = tiles * 16; break;
    case GS_HIP_BUF_TOTALS: p = v.totals, b = 64; break;
    case GS_HIP_BUF_TILE_ORDER: p = v.tile_order, b = tiles * 4; break;
    default: return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_buffer: unknown buffer");
  }
  *ptr = p;
  if (bytes != nullptr) *bytes = b;
  return 0;
}

int gs_hip_workspace_buffer(gs_hip_workspace* ws, int which, void** ptr, size_t* bytes) {
  return gs_hip_workspace_view_buffer(ws, 0, which, ptr, bytes);
}

int gs_hip_workspace_reserve_pairs(gs_hip_workspace* ws, uint64_t pairs) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "reserve_pairs: null workspace");
  for (gsh::View* v : ws->views) {
    const int rc = gsh::reserve_pairs(ws, *v, pairs);
    if (rc != 0) return rc;
  }
  return 0;
}

int gs_hip_workspace_load(gs_hip_workspace* ws, const float* params, const float* sh_rest, uint32_t count,
                          uint32_t sh_coeffs_in) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_load: null workspace");
  if (count > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_load: more slots than capacity");
  const size_t cap = ws->capacity, shn = cap * ws->sh_coeffs * 3;
  gsh::View& v0 = gsh::view0(ws);
  GS_HIP_TRY(hipDeviceSynchronize());
  int rc = 0;
  for (void* p : {(void*)ws->params, (void*)ws->m_params, (void*)ws->v_params}) {
    if (rc == 0) rc = zero(v0, p, cap * 64);
  }
  for (void* p : {(void*)ws->sh, (void*)ws->m_sh, (void*)ws->v_sh}) {
    if (rc == 0) rc = zero(v0, p, shn * 4);
  }
  if (rc == 0) rc = zero(v0, ws->stats, cap * 16);
  if (rc != 0) return rc;
  GS_HIP_TRY(hipStreamSynchronize(v0.stream));
  if (count == 0) return 0;
  GS_HIP_TRY(hipMemcpy(ws->params, params, size_t(count) * 64, hipMemcpyHostToDevice));
  if (ws->sh_coeffs > 0 && sh_rest != nullptr) {
    // [slot][coeff][rgb] -> planes [coeff][capacity][rgb]
    std::vector<float> planes(size_t(ws->sh_coeffs) * count * 3, 0.0f);
    const uint32_t used = std::min(sh_coeffs_in, ws->sh_coeffs);
    for (uint32_t i = 0; i < count; ++i) {
      for (uint32_t k = 0; k < used; ++k) {
        for (int c = 0; c < 3; ++c) {
          planes[(size_t(k) * count + i) * 3 + c] = sh_rest[(size_t(i) * sh_coeffs_in + k) * 3 + c];
        }
      }
    }
    for (uint32_t k = 0; k < ws->sh_coeffs; ++k) {
      GS_HIP_TRY(hipMemcpy(ws->sh + size_t(k) * cap * 3, planes.data() + size_t(k) * count * 3, size_t(count) * 12,
                           hipMemcpyHostToDevice));
    }
  }
  GS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int gs_hip_workspace_store(gs_hip_workspace* ws, float* params, float* sh_rest, uint32_t count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "workspace_store: null workspace");
  if (count > ws->capacity) return gsh::set_error(GS_HIP_E_CAPACITY, "workspace_store: more slots than capacity");
  GS_HIP_TRY(hipDeviceSynchronize());
  if (count == 0) return 0;
  if (params != nullptr) GS_HIP_TRY(hipMemcpy(params, ws->params, size_t(count) * 64, hipMemcpyDeviceToHost));
  if (sh_rest != nullptr && ws->sh_coeffs > 0) {
    std::vector<float> plane(size_t(count) * 3);
    for (uint32_t k = 0; k < ws->sh_coeffs; ++k) {
      GS_HIP_TRY(hipMemcpy(plane.data(), ws->sh + size_t(k) * ws->capacity * 3, plane.size() * 4,
                           hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < count; ++i) {
        for (int c = 0; c < 3; ++c) sh_rest[(size_t(i) * ws->sh_coeffs + k) * 3 + c] = plane[size_t(i) * 3 + c];
      }
    }
  }
  return 0;
}

int gs_hip_set_target(gs_hip_workspace* ws, const uint8_t* rgba8) {
  if (ws == nullptr || rgba8 == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_target: null");
  gsh::View& v = gsh::view0(ws);
  GS_HIP_TRY(hipMemcpyAsync(v.target, rgba8, size_t(ws->width) * ws->height * 4, hipMemcpyHostToDevice, v.stream));
  GS_HIP_TRY(hipStreamSynchronize(v.stream));
  return 0;
}

int gs_hip_set_target_device(gs_hip_workspace* ws, const void* device_rgba8) {
  if (ws == nullptr || device_rgba8 == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "set_target_device: null");
  gsh::View& v = gsh::view0(ws);
  GS_HIP_TRY(hipMemcpyAsync(v.target, device_rgba8, size_t(ws->width) * ws->height * 4, hipMemcpyDeviceToDevice,
                            v.stream));
  return 0;
}

int gs_hip_workspace_timings(gs_hip_workspace* ws, int enable, const char** names, double* ms, int capacity,
                             int* count) {
  if (ws == nullptr) return gsh::set_error(GS_HIP_E_ARGUMENT, "timings: null workspace");
  gsh::Timer& t = ws->timer;
  int n = 0;
  if (t.enabled && t.used > 0) {
    GS_HIP_TRY(hipDeviceSynchronize());
    t.ms.assign(t.used, 0.0);
    for (size_t i = 0; i < t.used; ++i) {
      float v = 0.0f;
      if (hipEventElapsedTime(&v, t.start[i], t.stop[i]) == hipSuccess) t.ms[i] = v;
    }
    n = int(std::min<size_t>(t.used, size_t(std::max(capacity, 0))));
    for (int i = 0; i < n; ++i) {
      if (names != nullptr) names[i] = t.names[i].c_str();
      if (ms != nullptr) ms[i] = t.ms[i];
    }
  }
  if (count != nullptr) *count = n;
  t.used = 0;
  t.enabled = enable != 0;
  return 0;
}

}  // extern "C"
