// Posed images a trainer fits (engine/include/gs/dataset.h:12-85): frames with
// an OpenGL camera-to-world pose and an RGBA8 image at training resolution.
// The transforms.json loader is not part of this build; datasets come from
// code (make_orbit_dataset) or from the caller.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gs/camera.h"

namespace gs {

struct DatasetFrame {
  std::string name;
  float c2w[16];              // camera-to-world, OpenGL convention, row-major
  std::vector<uint8_t> rgba;  // training resolution; alpha = subject mask (255 without masks)
  bool holdout = false;       // reserved for evaluation
};

struct Dataset {
  std::string name;
  uint32_t width = 0, height = 0;
  float fx = 0, fy = 0, cx = 0, cy = 0;
  bool has_masks = false;
  std::vector<DatasetFrame> frames;
  float camera_center[3] = {0, 0, 0};
  float camera_extent = 1.0f;  // 1.1 x the largest camera distance from camera_center

  size_t train_count() const;
  size_t holdout_count() const;
};

// Camera of a frame: the OpenGL c2w turned into scene -> camera OpenCV axes.
Camera dataset_camera(const Dataset& dataset, const DatasetFrame& frame);
// Sets camera_center (mean camera position) and camera_extent.
void compute_camera_bounds(Dataset& dataset);
// `views` cameras on a circle of `radius` at `height` around the origin,
// looking at it, vertical field of view `fov_y_deg`; images left black.
Dataset make_orbit_dataset(uint32_t views, uint32_t width, uint32_t height, float radius, float height_y,
                           float fov_y_deg, uint32_t holdout_every = 0);

}  // namespace gs
