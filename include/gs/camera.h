// Pinhole camera (engine/include/gs/camera.h:10-16): scene -> camera in
// OpenCV axes (x right, y down, z forward).
#pragma once

#include <cstdint>

#include "gs/splat_scene.h"

namespace gs {

// x_cam = rotation (row-major) * x_scene + translation; `position` is where
// the camera sits in the scene (-rotation^T translation); intrinsics in pixels.
struct Camera {
  float rotation[9];
  float translation[3];
  float position[3];
  float fx, fy, cx, cy;
  uint32_t width, height;
};

// A camera at `eye` looking at `target` with world up `up` and a vertical
// field of view of `fov_y_deg` degrees; principal point at the image centre.
Camera look_at(const float eye[3], const float target[3], const float up[3], uint32_t width, uint32_t height,
               float fov_y_deg);

// Orbit camera around a target in a Y-up display space (engine/src/camera.cpp:83-107).
// Angles in radians; a positive pitch looks down on the target.
class OrbitCamera {
 public:
  void reset(const SceneBounds& bounds, bool flip_y);
  // Drag by (dx, dy) pixels; zoom by `factor` (above 1: closer).
  void orbit(float dx_px, float dy_px);
  void zoom(float factor);
  Camera camera(uint32_t width, uint32_t height) const;

  float yaw = 0.0f;
  float pitch = 0.0f;
  float distance = 3.0f;
  float target[3] = {0.0f, 0.0f, 0.0f};
  float fov_short_side_deg = 50.0f;
  bool flip_y = false;  // rotate Y-down data (COLMAP / .ply) 180 degrees about X
};

}  // namespace gs
