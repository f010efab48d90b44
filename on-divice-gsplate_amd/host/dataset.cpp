// The transforms.json loader and the initial gaussians of a dataset (the role
// of engine/src/dataset.cpp:85-244, 320-525).  Frames decode in parallel on
// host threads; a frame whose files cannot be read is dropped, not fatal.
#include "gs/dataset.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "gs/image_io.h"
#include "gs/json.h"

namespace gs {
namespace {

constexpr float kShC0 = 0.28209479177387814f;

std::string join_path(const std::string& dir, const std::string& rel) {
  if (rel.empty() || rel[0] == '/') return rel;
  return dir + "/" + rel;
}

// Runs body(i) for i < n on up to 8 host threads; rethrows the first exception.
template <typename F>
void parallel_for(size_t n, F&& body) {
  const unsigned threads = std::max(1u, std::min(std::thread::hardware_concurrency(), 8u));
  std::atomic<size_t> next{0};
  std::exception_ptr error;
  std::mutex lock;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < std::min<size_t>(threads, n); ++t) {
    pool.emplace_back([&] {
      for (size_t i = next++; i < n; i = next++) {
        try {
          body(i);
        } catch (...) {
          std::lock_guard<std::mutex> g(lock);
          if (!error) error = std::current_exception();
        }
      }
    });
  }
  for (std::thread& t : pool) t.join();
  if (error) std::rethrow_exception(error);
}

// An OpenCV camera-space point through an OpenGL camera-to-world matrix
// (OpenCV (x, y, z) is OpenGL (x, -y, -z)).
void camera_to_world(const float c2w[16], float x, float y, float z, float out[3]) {
  for (int r = 0; r < 3; ++r) out[r] = c2w[r * 4] * x - c2w[r * 4 + 1] * y - c2w[r * 4 + 2] * z + c2w[r * 4 + 3];
}

// 21 bits per axis of a cell index.
uint64_t cell_key(int64_t x, int64_t y, int64_t z) {
  return (uint64_t(x & 0x1fffff) << 42) | (uint64_t(y & 0x1fffff) << 21) | uint64_t(z & 0x1fffff);
}

// 10 bits spread to every third bit (Morton interleave).
uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

}  // namespace

Dataset load_dataset(const std::string& directory, uint32_t target_width) {
  const std::vector<uint8_t> text = read_file(join_path(directory, "transforms.json"));
  const JsonValue root = parse_json(std::string_view(reinterpret_cast<const char*>(text.data()), text.size()));
  const JsonValue* frames = root.find("frames");
  if (frames == nullptr || frames->type != JsonValue::Type::Array || frames->array.empty()) {
    throw std::runtime_error("transforms.json: no frames");
  }
  const double src_w = root.number_or("w", 0), src_h = root.number_or("h", 0);
  const double fl_x = root.number_or("fl_x", 0), fl_y = root.number_or("fl_y", fl_x);
  if (src_w <= 0 || src_h <= 0 || fl_x <= 0) throw std::runtime_error("transforms.json: w, h or fl_x missing");
  if (target_width == 0) throw std::invalid_argument("load_dataset: target width 0");
  const double depth_scale = root.number_or("depth_unit_scale_factor", 0.001);

  Dataset ds;
  if (const JsonValue* lab = root.find("gsplat_lab")) {
    const JsonValue* up = lab->find("world_up");
    if (up != nullptr && up->type == JsonValue::Type::Array && up->array.size() == 3) {
      double len = 0.0;
      for (const JsonValue& v : up->array) len += v.number * v.number;
      len = std::sqrt(len);
      if (len > 0.0)
        for (int k = 0; k < 3; ++k) ds.world_up[k] = float(up->array[size_t(k)].number / len);
    }
  }
  const size_t slash = directory.find_last_of('/');
  ds.name = slash == std::string::npos ? directory : directory.substr(slash + 1);
  // The short side becomes target_width (portrait and landscape captures get the same pixel count).
  const double s = std::min(1.0, double(target_width) / std::min(src_w, src_h));
  ds.width = uint32_t(std::lround(src_w * s));
  ds.height = uint32_t(std::lround(src_h * s));
  ds.fx = float(fl_x * s);
  ds.fy = float(fl_y * s);
  ds.cx = float(root.number_or("cx", src_w / 2) * s);
  ds.cy = float(root.number_or("cy", src_h / 2) * s);

  const size_t n = frames->array.size();
  ds.frames.resize(n);
  std::vector<std::string> image(n), mask(n), depth(n), conf(n), errors(n);
  for (size_t i = 0; i < n; ++i) {
    const JsonValue& f = frames->array[i];
    DatasetFrame& frame = ds.frames[i];
    image[i] = f.string_or("file_path", "");
    if (image[i].empty()) throw std::runtime_error("transforms.json: a frame has no file_path");
    frame.name = image[i];
    const JsonValue* m = f.find("transform_matrix");
    if (m == nullptr || m->array.size() != 4) throw std::runtime_error("transforms.json: a frame has no transform_matrix");
    for (int r = 0; r < 4; ++r) {
      const JsonValue& row = m->array[size_t(r)];
      if (row.array.size() != 4) throw std::runtime_error("transforms.json: transform_matrix is not 4x4");
      for (int c = 0; c < 4; ++c) frame.c2w[r * 4 + c] = float(row.array[size_t(c)].number);
    }
    mask[i] = f.string_or("mask_path", "");
    depth[i] = f.string_or("depth_file_path", "");
    conf[i] = f.string_or("depth_confidence_path", "");
    frame.holdout = i % 8 == 0;
  }
  ds.has_masks = std::all_of(mask.begin(), mask.end(), [](const std::string& p) { return !p.empty(); });
  ds.has_depth = std::any_of(depth.begin(), depth.end(), [](const std::string& p) { return !p.empty(); });

  parallel_for(n, [&](size_t i) {
    DatasetFrame& frame = ds.frames[i];
    try {
      const Image8 img = decode_image(read_file(join_path(directory, image[i])), 4);
      if (img.width != uint32_t(src_w) || img.height != uint32_t(src_h)) {
        throw std::runtime_error("size differs from transforms.json");
      }
      frame.rgba = resize_area(img.data.data(), img.width, img.height, 4, ds.width, ds.height);
      const size_t pixels = size_t(ds.width) * ds.height;
      if (ds.has_masks) {
        const Image8 m = decode_image(read_file(join_path(directory, mask[i])), 1);
        const std::vector<uint8_t> a = resize_area(m.data.data(), m.width, m.height, 1, ds.width, ds.height);
        for (size_t p = 0; p < pixels; ++p) frame.rgba[p * 4 + 3] = a[p];
      } else {
        for (size_t p = 0; p < pixels; ++p) frame.rgba[p * 4 + 3] = 255;
      }
    } catch (const std::exception& e) {
      errors[i] = e.what();
      frame.rgba.clear();
      return;
    }
    if (depth[i].empty()) return;
    try {  // depth is optional per frame: an unreadable map just adds no points
      const Image16 d = decode_image16(read_file(join_path(directory, depth[i])));
      frame.depth.resize(d.data.size());
      for (size_t p = 0; p < d.data.size(); ++p) frame.depth[p] = float(d.data[p] * depth_scale);
      frame.depth_width = d.width;
      frame.depth_height = d.height;
      if (!conf[i].empty()) {
        Image8 c = decode_image(read_file(join_path(directory, conf[i])), 1);
        if (c.width == d.width && c.height == d.height) frame.depth_confidence = std::move(c.data);
      }
    } catch (const std::exception&) {
      frame.depth.clear();
      frame.depth_confidence.clear();
      frame.depth_width = frame.depth_height = 0;
    }
  });
  std::vector<DatasetFrame> kept;
  for (size_t i = 0; i < n; ++i) {
    if (errors[i].empty()) {
      kept.push_back(std::move(ds.frames[i]));
    } else {
      ds.skipped.push_back(image[i] + ": " + errors[i]);
    }
  }
  if (kept.size() < 2 || kept.size() * 2 < n) {
    throw std::runtime_error("too few readable frames (" + std::to_string(kept.size()) + " of " + std::to_string(n) +
                             ")" + (ds.skipped.empty() ? std::string() : ": " + ds.skipped.front()));
  }
  ds.frames = std::move(kept);

  const std::string ply = root.string_or("ply_file_path", "");
  if (!ply.empty()) {
    const std::vector<uint8_t> bytes = read_file(join_path(directory, ply));
    PointCloud pc = load_ply_points(bytes.data(), bytes.size());
    ds.points = std::move(pc.xyz);
    ds.point_colors = std::move(pc.rgb);
  }
  compute_camera_bounds(ds);
  return ds;
}

SplatScene initialize_gaussians(const Dataset& ds, const InitOptions& options, InitStats* stats) {
  std::vector<float> points, colors;  // colours 0..1
  std::string source;
  const bool use_mask = options.subject_only && ds.has_masks;
  if (ds.has_depth) {
    source = "depth";
    const uint8_t min_conf = uint8_t(std::clamp(options.min_confidence, 0.0f, 1.0f) * 255.0f);
    for (const DatasetFrame& f : ds.frames) {
      if (f.holdout || f.depth.empty()) continue;
      for (uint32_t v = 0; v < f.depth_height; ++v) {
        for (uint32_t u = 0; u < f.depth_width; ++u) {
          const size_t di = size_t(v) * f.depth_width + u;
          const float z = f.depth[di];
          if (z <= 0.0f) continue;
          if (!f.depth_confidence.empty() && f.depth_confidence[di] < min_conf) continue;
          // The depth sample's centre in training pixels, and the colour under it.
          const float px = (u + 0.5f) * ds.width / f.depth_width, py = (v + 0.5f) * ds.height / f.depth_height;
          const uint32_t ix = std::min(ds.width - 1, uint32_t(px)), iy = std::min(ds.height - 1, uint32_t(py));
          const uint8_t* rgba = &f.rgba[(size_t(iy) * ds.width + ix) * 4];
          if (use_mask && rgba[3] < 128) continue;
          float w[3];
          camera_to_world(f.c2w, (px - ds.cx) / ds.fx * z, (py - ds.cy) / ds.fy * z, z, w);
          points.insert(points.end(), w, w + 3);
          for (int c = 0; c < 3; ++c) colors.push_back(rgba[c] / 255.0f);
        }
      }
    }
  }
  if (points.empty() && !ds.points.empty()) {
    source = "points";
    points = ds.points;
    for (size_t i = 0; i < points.size(); ++i) colors.push_back(i < ds.point_colors.size() ? ds.point_colors[i] / 255.0f : 0.5f);
  }
  if (points.empty()) {
    source = "random";
    std::mt19937 rng(options.seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    const float radius = 0.5f * ds.camera_extent;
    for (uint32_t i = 0; i < options.max_gaussians; ++i) {
      float p[3];
      do {
        for (float& x : p) x = u(rng);
      } while (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > 1.0f);
      for (int k = 0; k < 3; ++k) points.push_back(ds.camera_center[k] + radius * p[k]);
      for (int k = 0; k < 3; ++k) colors.push_back(0.5f);
    }
  }
  const size_t raw = points.size() / 3;
  if (raw == 0) throw std::runtime_error("initialize_gaussians: no points");

  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < raw; ++i) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], points[i * 3 + k]);
      hi[k] = std::max(hi[k], points[i * 3 + k]);
    }
  }
  // The span of the points without the outer 0.5% per axis (depth smeared
  // off edges, flyers): it sizes the cell searches and caps the scales.
  float extent = 1e-6f;
  {
    std::vector<float> axis(raw);
    double e2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      for (size_t i = 0; i < raw; ++i) axis[i] = points[i * 3 + k];
      const size_t a = raw / 200, b = raw - 1 - raw / 200;
      std::nth_element(axis.begin(), axis.begin() + a, axis.end());
      const float low = axis[a];
      std::nth_element(axis.begin(), axis.begin() + b, axis.end());
      e2 += double(axis[b] - low) * (axis[b] - low);
    }
    extent += float(std::sqrt(e2));
  }
  auto cell_of = [&](const float* p, float cell, int64_t c[3]) {
    for (int k = 0; k < 3; ++k) c[k] = int64_t(std::floor((p[k] - lo[k]) / cell));
  };
  std::vector<uint64_t> keys;
  auto occupied = [&](const std::vector<float>& p, float cell) {
    keys.resize(p.size() / 3);
    for (size_t i = 0; i < keys.size(); ++i) {
      int64_t c[3];
      cell_of(&p[i * 3], cell, c);
      keys[i] = cell_key(c[0], c[1], c[2]);
    }
    std::sort(keys.begin(), keys.end());
    return size_t(std::unique(keys.begin(), keys.end()) - keys.begin());
  };
  // Cell size whose occupied-cell count falls just under `target` (12 bisections in log space).
  auto cell_for = [&](const std::vector<float>& p, size_t target) {
    float small = extent * 1e-4f, large = extent * 0.1f;
    for (int it = 0; it < 12; ++it) {
      const float mid = std::sqrt(small * large);
      (occupied(p, mid) > target ? small : large) = mid;
    }
    return large;
  };

  float voxel = 0.0f;
  std::vector<float> merged, merged_colors;
  if (raw <= options.max_gaussians) {
    merged = std::move(points);
    merged_colors = std::move(colors);
  } else {
    voxel = cell_for(points, options.max_gaussians);
    struct Acc {
      double p[3] = {0, 0, 0}, c[3] = {0, 0, 0};
      uint32_t n = 0;
    };
    std::map<uint64_t, Acc> cells;  // ordered: the merged points come out in key order
    for (size_t i = 0; i < raw; ++i) {
      int64_t c[3];
      cell_of(&points[i * 3], voxel, c);
      Acc& a = cells[cell_key(c[0], c[1], c[2])];
      for (int k = 0; k < 3; ++k) {
        a.p[k] += points[i * 3 + k];
        a.c[k] += colors[i * 3 + k];
      }
      ++a.n;
    }
    for (const auto& kv : cells) {
      for (int k = 0; k < 3; ++k) {
        merged.push_back(float(kv.second.p[k] / kv.second.n));
        merged_colors.push_back(float(kv.second.c[k] / kv.second.n));
      }
    }
  }
  const size_t n = merged.size() / 3;
  // Morton order (10 bits per axis over the bounds): neighbours in memory are neighbours in space.
  {
    std::vector<std::pair<uint32_t, uint32_t>> order(n);
    for (size_t i = 0; i < n; ++i) {
      uint32_t code = 0;
      for (int k = 0; k < 3; ++k) {
        const float span = std::max(hi[k] - lo[k], 1e-6f);
        code |= spread3(uint32_t(std::clamp((merged[i * 3 + k] - lo[k]) / span, 0.0f, 1.0f) * 1023.0f)) << k;
      }
      order[i] = {code, uint32_t(i)};
    }
    std::sort(order.begin(), order.end());
    std::vector<float> p2(merged.size()), c2(merged_colors.size());
    for (size_t j = 0; j < n; ++j) {
      std::copy_n(&merged[size_t(order[j].second) * 3], 3, &p2[j * 3]);
      std::copy_n(&merged_colors[size_t(order[j].second) * 3], 3, &c2[j * 3]);
    }
    merged = std::move(p2);
    merged_colors = std::move(c2);
  }
  // Scale: mean distance to the 3 nearest neighbours, searched over the 27 cells around each point.
  const float cell = 2.0f * (voxel > 0.0f ? voxel : cell_for(merged, std::max<size_t>(n / 2, 1)));
  std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
  for (uint32_t i = 0; i < n; ++i) {
    int64_t c[3];
    cell_of(&merged[size_t(i) * 3], cell, c);
    grid[cell_key(c[0], c[1], c[2])].push_back(i);
  }
  SplatScene scene;
  scene.convention = AxisConvention::RUB;  // dataset worlds are Y-up
  scene.resize(uint32_t(n), 0);
  const float logit = std::log(options.initial_opacity / (1.0f - options.initial_opacity));
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = &merged[size_t(i) * 3];
    int64_t c[3];
    cell_of(p, cell, c);
    float best[3] = {INFINITY, INFINITY, INFINITY};  // squared distances, ascending
    for (int dz = -1; dz <= 1; ++dz) {
      for (int dy = -1; dy <= 1; ++dy) {
        for (int dx = -1; dx <= 1; ++dx) {
          const auto it = grid.find(cell_key(c[0] + dx, c[1] + dy, c[2] + dz));
          if (it == grid.end()) continue;
          for (uint32_t j : it->second) {
            if (j == i) continue;
            const float* q = &merged[size_t(j) * 3];
            const float ex = p[0] - q[0], ey = p[1] - q[1], ez = p[2] - q[2];
            float d2 = ex * ex + ey * ey + ez * ez;
            for (int k = 0; k < 3; ++k) {
              if (d2 < best[k]) std::swap(d2, best[k]);
            }
          }
        }
      }
    }
    float sum = 0.0f;
    int found = 0;
    for (float d2 : best) {
      if (std::isfinite(d2)) {
        sum += std::sqrt(d2);
        ++found;
      }
    }
    const float dist = found > 0 ? sum / found : cell;
    const float scale = std::clamp(dist, 1e-4f * extent, 0.02f * extent);
    for (int k = 0; k < 3; ++k) {
      scene.positions[size_t(i) * 3 + k] = p[k];
      scene.log_scales[size_t(i) * 3 + k] = std::log(scale);
      scene.sh_dc[size_t(i) * 3 + k] = (merged_colors[size_t(i) * 3 + k] - 0.5f) / kShC0;
    }
    scene.rotations[size_t(i) * 4] = 1.0f;
    scene.opacity_logits[i] = logit;
  }
  if (stats != nullptr) {
    stats->source = source;
    stats->raw_points = raw;
    stats->voxel_size = voxel;
  }
  return scene;
}

}  // namespace gs
