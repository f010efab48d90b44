// Cameras and datasets of the host API (engine/src/camera.cpp,
// engine/src/dataset.cpp:66-83, 226-250 define the same conventions).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "gs/camera.h"
#include "gs/dataset.h"

namespace gs {
namespace {

// a x b: component i is a[i+1] b[i+2] - a[i+2] b[i+1], indices mod 3.
void cross(const double a[3], const double b[3], double out[3]) {
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    out[i] = a[j] * b[k] - a[k] * b[j];
  }
}

void normalise(double v[3]) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n > 0) v[0] /= n, v[1] /= n, v[2] /= n;
}

}  // namespace

Camera look_at(const float eye[3], const float target[3], const float up[3], uint32_t width, uint32_t height,
               float fov_y_deg) {
  double f[3] = {double(target[0]) - eye[0], double(target[1]) - eye[1], double(target[2]) - eye[2]};
  normalise(f);
  const double u[3] = {up[0], up[1], up[2]};
  double r[3], d[3];
  cross(f, u, r);
  normalise(r);
  cross(f, r, d);  // down = forward x right
  Camera c{};
  const double* rows[3] = {r, d, f};
  for (int i = 0; i < 3; ++i) {
    double t = 0;
    for (int k = 0; k < 3; ++k) {
      c.rotation[i * 3 + k] = float(rows[i][k]);
      t -= rows[i][k] * eye[k];
    }
    c.translation[i] = float(t);
    c.position[i] = eye[i];
  }
  c.fy = float((height / 2.0) / std::tan(fov_y_deg * 3.14159265358979 / 360.0));
  c.fx = c.fy;
  c.cx = width / 2.0f;
  c.cy = height / 2.0f;
  c.width = width;
  c.height = height;
  return c;
}

void OrbitCamera::reset(const SceneBounds& b, bool flip) {
  flip_y = flip;
  for (int k = 0; k < 3; ++k) target[k] = b.center[k];
  if (flip_y) target[1] = -target[1], target[2] = -target[2];
  distance = 2.5f * b.radius;
  yaw = 0.0f;
  pitch = 0.2f;
}

void OrbitCamera::orbit(float dx_px, float dy_px) {
  yaw += dx_px * 0.005f;
  pitch = std::clamp(pitch + dy_px * 0.005f, -1.5f, 1.5f);
}

void OrbitCamera::zoom(float factor) { distance = std::max(distance / std::max(factor, 1e-3f), 1e-3f); }

Camera OrbitCamera::camera(uint32_t width, uint32_t height) const {
  // Display space is Y-up; the eye sits `distance` from the target, yaw about
  // +Y, pitch raising it.  A flipped scene is turned half about X first.
  const float ce = std::cos(pitch), se = std::sin(pitch);
  const float eye_d[3] = {target[0] + distance * ce * std::sin(yaw), target[1] + distance * se,
                          target[2] + distance * ce * std::cos(yaw)};
  const float up[3] = {0.0f, 1.0f, 0.0f};
  const float short_side = float(std::min(width, height));
  const float fov_y = 2.0f * std::atan(std::tan(fov_short_side_deg * 3.14159265f / 360.0f) * height / short_side) *
                      180.0f / 3.14159265f;
  Camera c = look_at(eye_d, target, up, width, height, fov_y);
  if (flip_y) {
    // scene -> display is diag(1, -1, -1): fold it into the rotation.
    for (int r = 0; r < 3; ++r) c.rotation[r * 3 + 1] = -c.rotation[r * 3 + 1], c.rotation[r * 3 + 2] = -c.rotation[r * 3 + 2];
    c.position[1] = -c.position[1];
    c.position[2] = -c.position[2];
  }
  return c;
}

size_t Dataset::train_count() const {
  return size_t(std::count_if(frames.begin(), frames.end(), [](const DatasetFrame& f) { return !f.holdout; }));
}

size_t Dataset::holdout_count() const { return frames.size() - train_count(); }

Camera dataset_camera(const Dataset& ds, const DatasetFrame& frame) {
  // c2w (OpenGL: x right, y up, z back) -> scene -> camera with OpenCV axes:
  // R = diag(1, -1, -1) R_c2w^T, t = -R position.
  Camera c{};
  const float* m = frame.c2w;
  const float flip[3] = {1.0f, -1.0f, -1.0f};
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) c.rotation[r * 3 + k] = flip[r] * m[k * 4 + r];
    c.position[r] = m[r * 4 + 3];
  }
  for (int r = 0; r < 3; ++r) {
    c.translation[r] = -(c.rotation[r * 3] * c.position[0] + c.rotation[r * 3 + 1] * c.position[1] +
                         c.rotation[r * 3 + 2] * c.position[2]);
  }
  c.fx = ds.fx;
  c.fy = ds.fy;
  c.cx = ds.cx;
  c.cy = ds.cy;
  c.width = ds.width;
  c.height = ds.height;
  return c;
}

void compute_camera_bounds(Dataset& ds) {
  if (ds.frames.empty()) return;
  double mean[3] = {0, 0, 0};
  for (const DatasetFrame& f : ds.frames)
    for (int k = 0; k < 3; ++k) mean[k] += f.c2w[k * 4 + 3];
  for (double& v : mean) v /= double(ds.frames.size());
  double far = 0;
  for (const DatasetFrame& f : ds.frames) {
    double d = 0;
    for (int k = 0; k < 3; ++k) d += (f.c2w[k * 4 + 3] - mean[k]) * (f.c2w[k * 4 + 3] - mean[k]);
    far = std::max(far, std::sqrt(d));
  }
  for (int k = 0; k < 3; ++k) ds.camera_center[k] = float(mean[k]);
  ds.camera_extent = std::max(float(1.1 * far), 1e-3f);
}

Dataset make_orbit_dataset(uint32_t views, uint32_t width, uint32_t height, float radius, float height_y,
                           float fov_y_deg, uint32_t holdout_every) {
  if (views == 0 || width == 0 || height == 0) throw std::invalid_argument("make_orbit_dataset: empty");
  Dataset ds;
  ds.name = "orbit";
  ds.width = width;
  ds.height = height;
  ds.fy = float((height / 2.0) / std::tan(fov_y_deg * 3.14159265358979 / 360.0));
  ds.fx = ds.fy;
  ds.cx = width / 2.0f;
  ds.cy = height / 2.0f;
  for (uint32_t k = 0; k < views; ++k) {
    const double th = 2.0 * 3.14159265358979 * k / views;
    const float eye[3] = {float(radius * std::sin(th)), height_y, float(radius * std::cos(th))};
    const float origin[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    const Camera c = look_at(eye, origin, up, width, height, fov_y_deg);
    DatasetFrame f;
    f.name = "view_" + std::to_string(k);
    // c2w (OpenGL): columns right, up, back; the camera rows are right, down, forward.
    for (int r = 0; r < 3; ++r) {
      f.c2w[r * 4 + 0] = c.rotation[0 * 3 + r];
      f.c2w[r * 4 + 1] = -c.rotation[1 * 3 + r];
      f.c2w[r * 4 + 2] = -c.rotation[2 * 3 + r];
      f.c2w[r * 4 + 3] = eye[r];
    }
    f.c2w[12] = f.c2w[13] = f.c2w[14] = 0.0f;
    f.c2w[15] = 1.0f;
    f.rgba.assign(size_t(width) * height * 4, 0);
    for (size_t p = 0; p < size_t(width) * height; ++p) f.rgba[p * 4 + 3] = 255;
    f.holdout = holdout_every > 0 && views > 1 && k % holdout_every == 0;
    ds.frames.push_back(std::move(f));
  }
  compute_camera_bounds(ds);
  return ds;
}

}  // namespace gs
