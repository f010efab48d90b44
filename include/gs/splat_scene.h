// Gaussian splat scenes and their file formats: the scene type and I/O of
// the reference engine (engine/include/gs/splat_scene.h:1-78), same names
// and meanings.  Loaders throw std::runtime_error with a readable message.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gs {

// Which way a scene's x, y, z axes point.  RDF: x right, y down, z forward
// (COLMAP; what 3DGS .ply files hold).  RUB: x right, y up, z back (three.js;
// the axes SPZ files are stored in).
enum class AxisConvention { RDF, RUB };

// Splats in the 3DGS training parameterisation: log scales, opacity logits,
// unnormalised (w, x, y, z) quaternions, raw SH coefficients.
struct SplatScene {
  uint32_t count = 0;
  int sh_degree = 0;         // 0..3
  bool antialiased = false;  // SPZ flag: render with the mip-splatting opacity compensation
  AxisConvention convention = AxisConvention::RDF;
  // Per splat, splat-major: xyz; log sx sy sz; quaternion w x y z; logit;
  // DC rgb; then the rest coefficients as [coefficient][rgb].
  std::vector<float> positions;
  std::vector<float> log_scales;
  std::vector<float> rotations;
  std::vector<float> opacity_logits;
  std::vector<float> sh_dc;
  std::vector<float> sh_rest;

  // `n` zeroed splats of SH degree `degree`.
  void resize(uint32_t n, int degree);
};

// SH coefficients beyond DC: 0, 3, 8, 15 for degrees 0..3.
int sh_rest_coeffs(int degree);

// Where a scene sits: the median position on each axis and the distance
// from there that 90% of the splats lie within.
struct SceneBounds {
  float center[3];
  float radius;
};
SceneBounds compute_bounds(const SplatScene& scene);

// Binary little-endian PLY with the standard 3DGS vertex properties (any
// scalar property type on input; f_rest stored channel-major).
SplatScene load_ply(const uint8_t* data, size_t size);
// SPZ v2/v3 (gzip); v4 (zstd) is rejected with a message naming v4.
SplatScene load_spz(const uint8_t* data, size_t size);
// Detects PLY or SPZ.
SplatScene load_splat_file(const uint8_t* data, size_t size);
std::vector<uint8_t> save_ply(const SplatScene& scene);
// The header save_ply writes for `count` vertices of SH degree `sh_degree`.
std::string ply_header(uint32_t count, int sh_degree);
// RDF <-> RUB (a half turn about x): positions, rotations and SH signs.
void convert_convention(SplatScene& scene, AxisConvention target);
// A torus knot of anisotropic splats with degree-1 SH, in RUB convention.
SplatScene make_synthetic_scene(uint32_t count, uint32_t seed = 1);

// xyz (and rgb, 128 where absent) of the vertices of a binary little-endian
// PLY point cloud.
struct PointCloud {
  std::vector<float> xyz;
  std::vector<uint8_t> rgb;
};
PointCloud load_ply_points(const uint8_t* data, size_t size);

// Reads a whole file; throws std::runtime_error if it cannot.
std::vector<uint8_t> read_file(const std::string& path);
void write_file(const std::string& path, const std::vector<uint8_t>& data);

}  // namespace gs
