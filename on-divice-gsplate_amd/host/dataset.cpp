// Datasets on disk and the first gaussians of a scene.
//
// load_dataset reads a nerfstudio-style transforms.json folder (the layout the
// reference's captures write; engine/src/dataset.cpp:85-244 defines which
// keys mean what).  It works in three stages: the manifest is read and
// checked into plain FrameSpec records; the frames' files are decoded by a
// few std::async workers, each taking every k-th frame, each frame's outcome
// recorded as pixels or an error message; and the readable frames are kept
// in manifest order.  A frame that cannot be read is listed in
// Dataset::skipped; a dataset with fewer than two readable frames, or fewer
// than half of them, is refused.
//
// initialize_gaussians (engine/src/dataset.cpp:320-525 defines the steps)
// collects points from the training frames' depth maps, else the sparse
// cloud, else a random ball around the cameras; merges them per voxel when
// there are more than the budget; orders them along a Morton curve; and gives
// each a scale from its three nearest neighbours.  The point store here is a
// struct of 3-vectors, cells are hashed into one 64-bit key, the voxel size
// comes from a multiplicative bracket refined by bisection, and neighbours
// are found through a cell-sorted index (counting sort by cell) with rings
// grown until three are in reach.
#include "gs/dataset.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <future>
#include <random>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "gs/image_io.h"
#include "gs/json.h"

namespace gs {
namespace {

using Vec3 = std::array<float, 3>;

constexpr float kDcBasis = 0.28209479177387814f;  // Y_0^0: colour = kDcBasis * dc + 0.5

std::string resolve(const std::filesystem::path& dir, const std::string& relative) {
  if (relative.empty()) return relative;
  const std::filesystem::path p(relative);
  return p.is_absolute() ? relative : (dir / p).string();
}

// ---- the manifest ---------------------------------------------------------------

struct FrameSpec {
  std::string image, mask, depth, confidence;
  float c2w[16];
};

struct Manifest {
  double width = 0, height = 0, fx = 0, fy = 0, cx = 0, cy = 0;
  double depth_unit = 0.001;  // metres per stored depth unit
  std::vector<FrameSpec> frames;
  std::string ply;
  bool has_up = false;
  Vec3 up{0.0f, 1.0f, 0.0f};
};

[[noreturn]] void bad_manifest(const std::string& what) { throw std::runtime_error("transforms.json: " + what); }

Manifest read_manifest(const JsonValue& root) {
  Manifest m;
  m.width = root.number_or("w", 0.0);
  m.height = root.number_or("h", 0.0);
  m.fx = root.number_or("fl_x", 0.0);
  m.fy = root.number_or("fl_y", m.fx);
  m.cx = root.number_or("cx", m.width * 0.5);
  m.cy = root.number_or("cy", m.height * 0.5);
  m.depth_unit = root.number_or("depth_unit_scale_factor", m.depth_unit);
  m.ply = root.string_or("ply_file_path", "");
  if (!(m.width > 0 && m.height > 0 && m.fx > 0)) bad_manifest("w, h or fl_x missing");
  // A capture's own up direction (gsplat_lab.world_up), kept as a unit vector.
  if (const JsonValue* lab = root.find("gsplat_lab"); lab != nullptr) {
    const JsonValue* up = lab->find("world_up");
    if (up != nullptr && up->array.size() == 3) {
      const double x = up->array[0].number, y = up->array[1].number, z = up->array[2].number;
      const double norm = std::sqrt(x * x + y * y + z * z);
      if (norm > 0.0) {
        m.up = {float(x / norm), float(y / norm), float(z / norm)};
        m.has_up = true;
      }
    }
  }
  const JsonValue* list = root.find("frames");
  if (list == nullptr || list->type != JsonValue::Type::Array || list->array.empty()) bad_manifest("no frames");
  for (const JsonValue& f : list->array) {
    FrameSpec spec;
    spec.image = f.string_or("file_path", "");
    if (spec.image.empty()) bad_manifest("a frame has no file_path");
    const JsonValue* rows = f.find("transform_matrix");
    const bool square = rows != nullptr && rows->array.size() == 4 &&
                        std::all_of(rows->array.begin(), rows->array.end(),
                                    [](const JsonValue& r) { return r.array.size() == 4; });
    if (!square) bad_manifest("frame " + spec.image + ": transform_matrix is not 4x4");
    for (size_t k = 0; k < 16; ++k) spec.c2w[k] = float(rows->array[k / 4].array[k % 4].number);
    spec.mask = f.string_or("mask_path", "");
    spec.depth = f.string_or("depth_file_path", "");
    spec.confidence = f.string_or("depth_confidence_path", "");
    m.frames.push_back(std::move(spec));
  }
  return m;
}

// ---- decoding ----------------------------------------------------------------------

struct Decoded {
  DatasetFrame frame;
  std::string error;  // empty: readable
};

// One frame's files at the dataset's resolution (the mask, if any, replaces
// alpha; otherwise alpha is 255).  Depth is optional per frame: a map that
// cannot be read just leaves the frame without one.
Decoded decode_frame(const std::filesystem::path& dir, const Manifest& m, const FrameSpec& spec, uint32_t w,
                     uint32_t h, bool masks) {
  Decoded out;
  out.frame.name = spec.image;
  std::copy_n(spec.c2w, 16, out.frame.c2w);
  try {
    const Image8 photo = decode_image(read_file(resolve(dir, spec.image)), 4);
    if (photo.width != uint32_t(m.width) || photo.height != uint32_t(m.height)) {
      throw std::runtime_error("image is " + std::to_string(photo.width) + "x" + std::to_string(photo.height) +
                               ", transforms.json says " + std::to_string(uint32_t(m.width)) + "x" +
                               std::to_string(uint32_t(m.height)));
    }
    out.frame.rgba = resize_area(photo.data.data(), photo.width, photo.height, 4, w, h);
    std::vector<uint8_t> alpha(size_t(w) * h, 255);
    if (masks) {
      const Image8 mask = decode_image(read_file(resolve(dir, spec.mask)), 1);
      alpha = resize_area(mask.data.data(), mask.width, mask.height, 1, w, h);
    }
    for (size_t p = 0; p < alpha.size(); ++p) out.frame.rgba[p * 4 + 3] = alpha[p];
  } catch (const std::exception& e) {
    out.error = e.what();
    out.frame.rgba.clear();
    return out;
  }
  if (spec.depth.empty()) return out;
  try {
    const Image16 depth = decode_image16(read_file(resolve(dir, spec.depth)));
    std::vector<float> metres(depth.data.size());
    std::transform(depth.data.begin(), depth.data.end(), metres.begin(),
                   [&](uint16_t d) { return float(double(d) * m.depth_unit); });
    std::vector<uint8_t> conf;
    if (!spec.confidence.empty()) {
      Image8 c = decode_image(read_file(resolve(dir, spec.confidence)), 1);
      if (c.width == depth.width && c.height == depth.height) conf = std::move(c.data);
    }
    out.frame.depth = std::move(metres);
    out.frame.depth_confidence = std::move(conf);
    out.frame.depth_width = depth.width;
    out.frame.depth_height = depth.height;
  } catch (const std::exception&) {
    // no depth for this frame
  }
  return out;
}

}  // namespace

Dataset load_dataset(const std::string& directory, uint32_t target_width) {
  if (target_width == 0) throw std::invalid_argument("load_dataset: target width 0");
  const std::filesystem::path dir(directory);
  const std::vector<uint8_t> text = read_file((dir / "transforms.json").string());
  const Manifest m = read_manifest(parse_json(std::string_view(reinterpret_cast<const char*>(text.data()), text.size())));

  Dataset ds;
  ds.name = dir.filename().empty() ? dir.parent_path().filename().string() : dir.filename().string();
  if (m.has_up) std::copy(m.up.begin(), m.up.end(), ds.world_up);
  // Training resolution: the short side brought to target_width, never enlarged;
  // the intrinsics scale with it.
  const double scale = std::min(1.0, double(target_width) / std::min(m.width, m.height));
  ds.width = uint32_t(std::lround(m.width * scale));
  ds.height = uint32_t(std::lround(m.height * scale));
  ds.fx = float(m.fx * scale), ds.fy = float(m.fy * scale);
  ds.cx = float(m.cx * scale), ds.cy = float(m.cy * scale);
  ds.has_masks = std::none_of(m.frames.begin(), m.frames.end(), [](const FrameSpec& f) { return f.mask.empty(); });
  ds.has_depth = std::any_of(m.frames.begin(), m.frames.end(), [](const FrameSpec& f) { return !f.depth.empty(); });

  // Decode: worker t takes frames t, t + k, t + 2k, ...
  const size_t total = m.frames.size();
  std::vector<Decoded> decoded(total);
  const size_t workers = std::clamp<size_t>(std::thread::hardware_concurrency(), 1, 8);
  std::vector<std::future<void>> jobs;
  for (size_t t = 0; t < std::min(workers, total); ++t) {
    jobs.push_back(std::async(std::launch::async, [&, t] {
      for (size_t i = t; i < total; i += workers)
        decoded[i] = decode_frame(dir, m, m.frames[i], ds.width, ds.height, ds.has_masks);
    }));
  }
  for (std::future<void>& j : jobs) j.get();

  // Every 8th frame of the manifest (from the first) is held out for evaluation.
  size_t readable = 0;
  for (size_t i = 0; i < total; ++i) {
    Decoded& d = decoded[i];
    if (!d.error.empty()) {
      ds.skipped.push_back(m.frames[i].image + ": " + d.error);
      continue;
    }
    d.frame.holdout = i % 8 == 0;
    ds.frames.push_back(std::move(d.frame));
    ++readable;
  }
  const bool enough = readable >= 2 && 2 * readable >= total;
  if (!enough) {
    std::string msg = "load_dataset: only " + std::to_string(readable) + " of " + std::to_string(total) +
                      " frames are readable";
    if (!ds.skipped.empty()) msg += " (first failure: " + ds.skipped.front() + ")";
    throw std::runtime_error(msg);
  }
  if (!m.ply.empty()) {
    const std::vector<uint8_t> bytes = read_file(resolve(dir, m.ply));
    PointCloud cloud = load_ply_points(bytes.data(), bytes.size());
    ds.points = std::move(cloud.xyz);
    ds.point_colors = std::move(cloud.rgb);
  }
  compute_camera_bounds(ds);
  return ds;
}

// ---- initial gaussians --------------------------------------------------------------

namespace {

struct Points {
  std::vector<Vec3> at;
  std::vector<Vec3> colour;  // 0..1
};

// Every confident depth sample of the training frames, lifted to the world
// through its frame's camera (OpenCV x right, y down, z forward are the
// OpenGL pose's x, -y, -z), with the colour of the training pixel under the
// sample's centre; with `subject_only` and masks, samples off the subject
// (alpha < 128) are dropped.
void points_from_depth(const Dataset& ds, const InitOptions& o, Points& pts) {
  const float min_conf = std::clamp(o.min_confidence, 0.0f, 1.0f) * 255.0f;
  const bool subject = o.subject_only && ds.has_masks;
  for (const DatasetFrame& f : ds.frames) {
    if (f.holdout || f.depth.empty()) continue;
    const float sx = float(ds.width) / float(f.depth_width), sy = float(ds.height) / float(f.depth_height);
    for (size_t s = 0; s < f.depth.size(); ++s) {
      const float z = f.depth[s];
      if (!(z > 0.0f)) continue;
      if (!f.depth_confidence.empty() && float(f.depth_confidence[s]) < float(uint8_t(min_conf))) continue;
      const uint32_t u = uint32_t(s % f.depth_width), v = uint32_t(s / f.depth_width);
      const float px = (float(u) + 0.5f) * sx, py = (float(v) + 0.5f) * sy;
      const size_t pixel = size_t(std::min(uint32_t(py), ds.height - 1)) * ds.width + std::min(uint32_t(px), ds.width - 1);
      const uint8_t* rgba = &f.rgba[pixel * 4];
      if (subject && rgba[3] < 128) continue;
      const float cam[3] = {(px - ds.cx) / ds.fx * z, -((py - ds.cy) / ds.fy * z), -z};  // OpenGL camera axes
      Vec3 w;
      for (int r = 0; r < 3; ++r) w[r] = f.c2w[r * 4] * cam[0] + f.c2w[r * 4 + 1] * cam[1] + f.c2w[r * 4 + 2] * cam[2] + f.c2w[r * 4 + 3];
      pts.at.push_back(w);
      pts.colour.push_back({rgba[0] / 255.0f, rgba[1] / 255.0f, rgba[2] / 255.0f});
    }
  }
}

// `count` points uniform in the ball of `radius` around `centre` (direction
// from three normal draws, radius from the cube root of a uniform draw), grey.
void points_in_ball(const float centre[3], float radius, uint32_t count, uint32_t seed, Points& pts) {
  std::mt19937_64 gen(seed);
  std::normal_distribution<float> normal(0.0f, 1.0f);
  std::uniform_real_distribution<float> unit(0.0f, 1.0f);
  while (pts.at.size() < count) {
    const float d[3] = {normal(gen), normal(gen), normal(gen)};
    const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(len > 1e-12f)) continue;
    const float r = radius * std::cbrt(unit(gen)) / len;
    pts.at.push_back({centre[0] + r * d[0], centre[1] + r * d[1], centre[2] + r * d[2]});
    pts.colour.push_back({0.5f, 0.5f, 0.5f});
  }
}

// Length of the diagonal of the box spanned by the points after the outer
// 0.5% per axis is set aside (depth smeared off edges, flyers), from the
// 0.5% and 99.5% order statistics of a sorted subsample of at most 65536
// points per axis.
float robust_extent(const std::vector<Vec3>& at) {
  const size_t n = at.size(), stride = std::max<size_t>(1, n / 65536);
  double sq = 0.0;
  std::vector<float> axis;
  for (int k = 0; k < 3; ++k) {
    axis.clear();
    for (size_t i = 0; i < n; i += stride) axis.push_back(at[i][k]);
    std::sort(axis.begin(), axis.end());
    const size_t m = axis.size(), lo = m / 200, hi = m - 1 - m / 200;
    sq += double(axis[hi] - axis[lo]) * double(axis[hi] - axis[lo]);
  }
  return float(std::sqrt(sq)) + 1e-6f;
}

// Cells of side `cell` over the points' lower corner, three signed 21-bit
// indices packed into one key.
struct Grid {
  Vec3 origin;
  float cell;
  std::array<int64_t, 3> index(const Vec3& p) const {
    return {int64_t(std::floor((p[0] - origin[0]) / cell)), int64_t(std::floor((p[1] - origin[1]) / cell)),
            int64_t(std::floor((p[2] - origin[2]) / cell))};
  }
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    const uint64_t m = (1ull << 21) - 1;
    return (uint64_t(x) & m) | ((uint64_t(y) & m) << 21) | ((uint64_t(z) & m) << 42);
  }
  uint64_t key(const Vec3& p) const {
    const auto c = index(p);
    return key(c[0], c[1], c[2]);
  }
};

size_t occupancy(const std::vector<Vec3>& at, const Grid& g) {
  std::unordered_set<uint64_t> cells;
  cells.reserve(at.size());
  for (const Vec3& p : at) cells.insert(g.key(p));
  return cells.size();
}

// The cell side (searched between 1e-4 and 0.1 of the extent) whose
// occupancy is at most `budget`, as small as the search resolves: a
// doubling bracket, then ten bisections of its logarithm.
float cell_for_budget(const std::vector<Vec3>& at, const Vec3& origin, float extent, size_t budget) {
  float hi = extent * 1e-4f;
  while (hi < extent * 0.1f && occupancy(at, Grid{origin, hi}) > budget) hi *= 2.0f;
  hi = std::min(hi, extent * 0.1f);
  float lo = std::max(hi * 0.5f, extent * 1e-4f);
  for (int it = 0; it < 10; ++it) {
    const float mid = std::sqrt(lo * hi);
    if (occupancy(at, Grid{origin, mid}) > budget) lo = mid;
    else hi = mid;
  }
  return hi;
}

// Averages the points (and colours) of each occupied cell; cells in key order.
void merge_cells(Points& pts, const Grid& g) {
  std::unordered_map<uint64_t, size_t> slot;
  std::vector<std::pair<uint64_t, std::array<double, 7>>> acc;
  for (size_t i = 0; i < pts.at.size(); ++i) {
    const uint64_t k = g.key(pts.at[i]);
    auto [it, fresh] = slot.try_emplace(k, acc.size());
    if (fresh) acc.push_back({k, {0, 0, 0, 0, 0, 0, 0}});
    std::array<double, 7>& a = acc[it->second].second;
    for (int c = 0; c < 3; ++c) {
      a[c] += pts.at[i][c];
      a[3 + c] += pts.colour[i][c];
    }
    a[6] += 1.0;
  }
  std::sort(acc.begin(), acc.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  pts.at.resize(acc.size());
  pts.colour.resize(acc.size());
  for (size_t i = 0; i < acc.size(); ++i) {
    const std::array<double, 7>& a = acc[i].second;
    for (int c = 0; c < 3; ++c) {
      pts.at[i][c] = float(a[c] / a[6]);
      pts.colour[i][c] = float(a[3 + c] / a[6]);
    }
  }
}

// Morton (Z-order) code of 10-bit coordinates: bit b of axis k goes to bit 3 b + k.
uint32_t morton3(const uint32_t q[3]) {
  uint32_t code = 0;
  for (int b = 0; b < 10; ++b)
    for (int k = 0; k < 3; ++k) code |= ((q[k] >> b) & 1u) << (3 * b + k);
  return code;
}

void morton_order(Points& pts, const Vec3& lo, const Vec3& hi) {
  const size_t n = pts.at.size();
  std::vector<std::pair<uint32_t, uint32_t>> keyed(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      const float t = (pts.at[i][k] - lo[k]) / std::max(hi[k] - lo[k], 1e-6f);
      q[k] = uint32_t(std::clamp(t, 0.0f, 1.0f) * 1023.0f);
    }
    keyed[i] = {morton3(q), uint32_t(i)};
  }
  std::sort(keyed.begin(), keyed.end());
  Points sorted;
  sorted.at.reserve(n);
  sorted.colour.reserve(n);
  for (const auto& [code, i] : keyed) {
    sorted.at.push_back(pts.at[i]);
    sorted.colour.push_back(pts.colour[i]);
  }
  pts = std::move(sorted);
}

// The mean distance from each point to its three nearest others.  Points are
// counting-sorted by cell into one index array; each point searches the
// cube of cells around its own, one ring wider at a time, until three
// others are found within the searched radius (or three rings were
// searched: then whatever it found, or `fallback` for none).
std::vector<float> neighbour_spacing(const std::vector<Vec3>& at, const Grid& g, float fallback) {
  const size_t n = at.size();
  std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> range;  // cell -> [first, end) in `by_cell`
  std::vector<uint64_t> keys(n);
  for (size_t i = 0; i < n; ++i) {
    keys[i] = g.key(at[i]);
    ++range[keys[i]].second;
  }
  uint32_t run = 0;
  for (auto& [k, r] : range) {
    r.first = run;
    run += r.second;
    r.second = r.first;  // fill cursor
  }
  std::vector<uint32_t> by_cell(n);
  for (size_t i = 0; i < n; ++i) by_cell[range[keys[i]].second++] = uint32_t(i);
  std::vector<float> out(n, fallback);
  for (size_t i = 0; i < n; ++i) {
    const auto c = g.index(at[i]);
    float best[3] = {INFINITY, INFINITY, INFINITY};  // squared, ascending
    for (int ring = 1; ring <= 3; ++ring) {
      for (int dz = -ring; dz <= ring; ++dz)
        for (int dy = -ring; dy <= ring; ++dy)
          for (int dx = -ring; dx <= ring; ++dx) {
            if (std::max({std::abs(dx), std::abs(dy), std::abs(dz)}) != ring && ring > 1) continue;  // the new shell
            const auto it = range.find(Grid::key(c[0] + dx, c[1] + dy, c[2] + dz));
            if (it == range.end()) continue;
            for (uint32_t k = it->second.first; k < it->second.second; ++k) {
              const uint32_t j = by_cell[k];
              if (j == i) continue;
              float d2 = 0.0f;
              for (int a = 0; a < 3; ++a) d2 += (at[i][a] - at[j][a]) * (at[i][a] - at[j][a]);
              for (float& b : best)
                if (d2 < b) std::swap(d2, b);
            }
          }
      // Points within `ring` cells of this one's cell are all seen; a third
      // neighbour closer than ring cells' width cannot be further out.
      if (best[2] <= (float(ring) * g.cell) * (float(ring) * g.cell)) break;
    }
    float sum = 0.0f;
    int found = 0;
    for (float b : best)
      if (std::isfinite(b)) sum += std::sqrt(b), ++found;
    if (found > 0) out[i] = sum / float(found);
  }
  return out;
}

}  // namespace

SplatScene initialize_gaussians(const Dataset& ds, const InitOptions& options, InitStats* stats) {
  Points pts;
  std::string source = "random";
  if (ds.has_depth) points_from_depth(ds, options, pts);
  if (!pts.at.empty()) {
    source = "depth";
  } else if (!ds.points.empty()) {
    source = "points";
    for (size_t i = 0; i + 2 < ds.points.size(); i += 3) {
      pts.at.push_back({ds.points[i], ds.points[i + 1], ds.points[i + 2]});
      const auto col = [&](size_t c) { return i + c < ds.point_colors.size() ? ds.point_colors[i + c] / 255.0f : 0.5f; };
      pts.colour.push_back({col(0), col(1), col(2)});
    }
  } else {
    points_in_ball(ds.camera_center, 0.5f * ds.camera_extent, options.max_gaussians, options.seed, pts);
  }
  const size_t raw = pts.at.size();
  if (raw == 0) throw std::runtime_error("initialize_gaussians: no points");

  Vec3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const Vec3& p : pts.at)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], p[k]), hi[k] = std::max(hi[k], p[k]);
  const float extent = robust_extent(pts.at);

  float voxel = 0.0f;
  if (raw > options.max_gaussians) {
    voxel = cell_for_budget(pts.at, lo, extent, options.max_gaussians);
    merge_cells(pts, Grid{lo, voxel});
  }
  morton_order(pts, lo, hi);
  // Neighbour search cells: twice the voxel, or (unmerged) twice the cell
  // that would halve the points.
  const size_t n = pts.at.size();
  const float cell = 2.0f * (voxel > 0.0f ? voxel : cell_for_budget(pts.at, lo, extent, std::max<size_t>(n / 2, 1)));
  const std::vector<float> spacing = neighbour_spacing(pts.at, Grid{lo, cell}, cell);

  SplatScene scene;
  scene.convention = AxisConvention::RUB;  // dataset worlds are Y-up
  scene.resize(uint32_t(n), 0);
  const float logit = std::log(options.initial_opacity / (1.0f - options.initial_opacity));
  for (size_t i = 0; i < n; ++i) {
    const float log_scale = std::log(std::clamp(spacing[i], 1e-4f * extent, 0.02f * extent));
    for (int k = 0; k < 3; ++k) {
      scene.positions[i * 3 + k] = pts.at[i][k];
      scene.log_scales[i * 3 + k] = log_scale;
      scene.sh_dc[i * 3 + k] = (pts.colour[i][k] - 0.5f) / kDcBasis;
    }
    scene.rotations[i * 4] = 1.0f;
    scene.opacity_logits[i] = logit;
  }
  if (stats != nullptr) *stats = InitStats{source, raw, voxel};
  return scene;
}

}  // namespace gs
