// Drop-in BundleAdjustmentProblem for the MI355X SBA: the same namespace, type names, member names,
// defaults and memory layout as cuVSLAM's libs/sba/bundle_adjustment_problem.h:34-68 and
// libs/camera/rig.h:29-43, so a caller of SchurComplementBundlerGpu::solve compiles unchanged.
//
// When the including translation unit already uses Eigen (CUVSLAM_SBA_USE_EIGEN defined, as in the
// reference tree) the reference's own Eigen types are used. Otherwise small layout-identical PODs stand in
// for Vector2f/Vector3f (packed floats), Matrix2f and Isometry3f (column-major 2x2 and 4x4 floats).
#pragma once

#include <cstdint>
#include <vector>

#ifdef CUVSLAM_SBA_USE_EIGEN
#include <Eigen/Core>
#include <Eigen/Geometry>
namespace cuvslam {
using Vector2T = Eigen::Vector2f;
using Vector3T = Eigen::Vector3f;
using Matrix2T = Eigen::Matrix2f;
using Isometry3T = Eigen::Transform<float, 3, Eigen::Isometry>;
}  // namespace cuvslam
#else
namespace cuvslam {

struct Vector2T {
  float v[2] = {0.f, 0.f};
  Vector2T() = default;
  Vector2T(float x, float y) : v{x, y} {}
  float& x() { return v[0]; }
  float& y() { return v[1]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
  float* data() { return v; }
  const float* data() const { return v; }
};

struct Vector3T {
  float v[3] = {0.f, 0.f, 0.f};
  Vector3T() = default;
  Vector3T(float x, float y, float z) : v{x, y, z} {}
  float& x() { return v[0]; }
  float& y() { return v[1]; }
  float& z() { return v[2]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float z() const { return v[2]; }
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
  float* data() { return v; }
  const float* data() const { return v; }
};

struct Matrix2T {  // column-major
  float m[4] = {0.f, 0.f, 0.f, 0.f};
  static Matrix2T Identity() {
    Matrix2T r;
    r.m[0] = r.m[3] = 1.f;
    return r;
  }
  float& operator()(int r, int c) { return m[c * 2 + r]; }
  float operator()(int r, int c) const { return m[c * 2 + r]; }
  float* data() { return m; }
  const float* data() const { return m; }
};

struct alignas(16) Isometry3T {  // column-major 4x4, like Eigen::Transform<float, 3, Isometry>
  float m[16] = {1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f};
  static Isometry3T Identity() { return Isometry3T(); }
  float& operator()(int r, int c) { return m[c * 4 + r]; }
  float operator()(int r, int c) const { return m[c * 4 + r]; }
  float* data() { return m; }
  const float* data() const { return m; }
};

}  // namespace cuvslam
#endif

namespace cuvslam::camera {

struct ICameraModel;

// libs/camera/rig.h:29-43
struct Rig {
  static constexpr const int32_t kMaxCameras{32};
  Isometry3T camera_from_rig[kMaxCameras];
  const ICameraModel* intrinsics[kMaxCameras]{};  // not used by SBA
  int32_t num_cameras{0};
};

}  // namespace cuvslam::camera

namespace cuvslam::sba {

// sorting (as the reference documents it):
// - all observations are ordered by key frame
// - within each key frame observations are ordered by point id
struct BundleAdjustmentProblem {
  // last num_fixed_points points are not allowed to move
  std::vector<Vector3T> points;
  int32_t num_fixed_points = 0;

  std::vector<Matrix2T> info_matrix;  // unused, kept for layout parity
  std::vector<Vector2T> observation_xys;
  std::vector<Matrix2T> observation_infos;

  std::vector<int32_t> point_ids;
  std::vector<int32_t> pose_ids;
  std::vector<int8_t> camera_ids;

  // transformation from world/map space to rig space; the last num_fixed_key_frames do not move
  std::vector<Isometry3T> rig_from_world;
  int32_t num_fixed_key_frames = 0;

  camera::Rig rig;

  // solver options
  int32_t max_iterations = 20;
  float robustifier_scale = 1.f;

  // solver statistic
  int32_t iterations = 0;
  float initial_cost = 0;
  float last_cost = 0;
};

}  // namespace cuvslam::sba
