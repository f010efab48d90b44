// Posed images a trainer fits (engine/include/gs/dataset.h:12-85): frames with
// an OpenGL camera-to-world pose and an RGBA8 image at training resolution,
// from a nerfstudio transforms.json folder (load_dataset), from code
// (make_orbit_dataset) or from the caller; and the initial gaussians built
// from a dataset's depth maps, sparse points or cameras.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gs/camera.h"
#include "gs/splat_scene.h"

namespace gs {

// One photo of a capture.  c2w: camera-to-world, row-major, OpenGL axes
// (x right, y up, looking down -z).  rgba: the photo at the training
// resolution, alpha the subject mask (255 where there is none).  depth:
// metres at its own (usually lower) resolution, 0 where unknown, with an
// optional 0..255 confidence per sample.  holdout: kept for evaluation.
struct DatasetFrame {
  std::string name;
  float c2w[16];
  std::vector<uint8_t> rgba;
  uint32_t depth_width = 0, depth_height = 0;
  std::vector<float> depth;
  std::vector<uint8_t> depth_confidence;
  bool holdout = false;
};

// A capture at one training resolution with shared pinhole intrinsics.
// points / point_colors: a sparse cloud (xyz, rgb) when the capture has one.
// camera_center / camera_extent: the mean camera position and 1.1 times the
// farthest camera from it (the scene's length scale).  world_up: the
// capture's up direction when it states one.  skipped: frames whose files
// could not be read.
struct Dataset {
  std::string name;
  uint32_t width = 0, height = 0;
  float fx = 0, fy = 0, cx = 0, cy = 0;
  bool has_masks = false;
  bool has_depth = false;
  std::vector<DatasetFrame> frames;
  std::vector<float> points;
  std::vector<uint8_t> point_colors;
  float camera_center[3] = {0, 0, 0};
  float camera_extent = 1.0f;
  float world_up[3] = {0, 1, 0};
  std::vector<std::string> skipped;

  // Frames trained on / held out.
  size_t train_count() const;
  size_t holdout_count() const;
};

// Loads `directory`/transforms.json (nerfstudio layout: w, h, fl_x[, fl_y, cx,
// cy], frames with file_path and transform_matrix, optional mask_path,
// depth_file_path, depth_confidence_path, depth_unit_scale_factor,
// ply_file_path).  Images are area-resampled so their short side is
// `target_width` (never enlarged); every 8th frame is held out.  A frame
// whose files cannot be read is skipped and listed in `skipped`; fewer than
// two readable frames, or fewer than half, throws std::runtime_error.
Dataset load_dataset(const std::string& directory, uint32_t target_width);

// Camera of a frame: the OpenGL c2w turned into scene -> camera OpenCV axes.
Camera dataset_camera(const Dataset& dataset, const DatasetFrame& frame);
// Sets camera_center (mean camera position) and camera_extent.
void compute_camera_bounds(Dataset& dataset);
// `views` cameras on a circle of `radius` at `height` around the origin,
// looking at it, vertical field of view `fov_y_deg`; images left black.
Dataset make_orbit_dataset(uint32_t views, uint32_t width, uint32_t height, float radius, float height_y,
                           float fov_y_deg, uint32_t holdout_every = 0);

// How the initial gaussians are built: at most max_gaussians of them; with
// masks and subject_only, depth samples outside the subject are dropped, as
// are samples whose confidence (scaled to 0..1) is below min_confidence.
struct InitOptions {
  uint32_t max_gaussians = 150000;
  bool subject_only = true;
  float min_confidence = 0.3f;
  float initial_opacity = 0.2f;
  uint32_t seed = 1;
};

// What initialisation used: "depth", "points" or "random"; the points it
// started from; the voxel size they were merged at (0: no merging).
struct InitStats {
  std::string source;
  size_t raw_points = 0;
  float voxel_size = 0;
};

// Initial gaussians (SH degree 0, RUB) from the training frames' depth maps,
// else the sparse points, else random points in a ball of half the camera
// extent around the cameras' centre.  Over budget, points are merged per
// voxel (the cell size whose occupancy falls just under max_gaussians);
// gaussians are laid out in Morton order; each one's scale is the mean
// distance to its three nearest neighbours, clamped to [1e-4, 2e-2] of the
// points' extent.
SplatScene initialize_gaussians(const Dataset& dataset, const InitOptions& options, InitStats* stats = nullptr);

}  // namespace gs
