// Drop-in inertial SBA types for the MI355X path: the namespaces, type names, member names and defaults of
// cuVSLAM's common/imu_calibration.h:29-60, common/imu_measurement.h, libs/imu/imu_preintegration.h:33-108,
// libs/imu/imu_sba_problem.h:30-91 and libs/imu/imu_sba_gpu.h:42-47, so a caller of
// IMUBundlerGpuFixedVel::solve compiles unchanged. Matrices are column-major like Eigen's. The preintegration is
// computed by the library (sba_imu_integrate: FP64 integration of the samples, stored in FP32); GetDelta* and the
// information matrices follow imu_preintegration.cpp:151-240.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "cuvslam_sba_imu.h"
#include "sba/bundle_adjustment_problem.h"
#include "sba/small_matrix.h"

namespace cuvslam::imu {

struct ImuMeasurement {  // common/imu_measurement.h
  int64_t time_ns = 0;
  Vector3T linear_acceleration;
  Vector3T angular_velocity;
};

class ImuCalibration {  // common/imu_calibration.h:29-60 (EuRoC defaults)
public:
  ImuCalibration() = default;
  ImuCalibration(const Isometry3T& rig_from_imu, float gyroscope_noise_density, float gyroscope_random_walk,
                 float accelerometer_noise_density, float accelerometer_random_walk, float frequency)
      : rig_from_imu_(rig_from_imu), gnd_(gyroscope_noise_density), grw_(gyroscope_random_walk),
        and_(accelerometer_noise_density), arw_(accelerometer_random_walk), freq_(frequency) {}
  const Isometry3T& rig_from_imu() const { return rig_from_imu_; }
  float frequency() const { return freq_; }
  sba_imu_calibration view() const {
    sba_imu_calibration c{};
    std::memcpy(c.rig_from_imu, rig_from_imu_.data(), sizeof(c.rig_from_imu));
    c.gyroscope_noise_density = gnd_;
    c.gyroscope_random_walk = grw_;
    c.accelerometer_noise_density = and_;
    c.accelerometer_random_walk = arw_;
    c.frequency = freq_;
    return c;
  }

private:
  Isometry3T rig_from_imu_ = Isometry3T::Identity();
  float gnd_ = 0.00016968f;
  float grw_ = 0.000019393f;
  float and_ = 0.002f;
  float arw_ = 0.003f;
  float freq_ = 200.f;
};

}  // namespace cuvslam::imu

namespace cuvslam::sba_imu {

const float GRAVITY_VALUE = 9.81f;

class IMUPreintegration {
public:
  // covariance for [rotation, velocity, translation]
  Matrix9T cov_matrix_ = Matrix9T::Zero();
  Matrix3T acc_random_walk_accum_cov_matrix_ = Matrix3T::Zero();
  Matrix3T gyro_random_walk_accum_cov_matrix_ = Matrix3T::Zero();
  Matrix3T dR = Matrix3T::Identity();
  Vector3T dV;
  Vector3T dP;
  Matrix3T JRg, JVg, JVa, JPg, JPa;
  float dT_s = 0.f;

  IMUPreintegration() = default;
  IMUPreintegration(const Vector3T& gyro_bias, const Vector3T& acc_bias) { Initialize(gyro_bias, acc_bias); }

  // The empty integration at the given biases; the samples stay (imu_preintegration.cpp:40-67).
  void Initialize(const Vector3T& gyro_bias, const Vector3T& acc_bias) {
    const sba_imu_calibration none = imu::ImuCalibration().view();
    sba_imu_preintegration s{};
    sba_imu_integrate(&none, nullptr, 0, gyro_bias.data(), acc_bias.data(), &s);
    load(s);
  }
  // The state is a function of the samples and the integration biases: every call integrates all samples again
  // (a key frame holds a few dozen), which equals the reference's incremental update of a consistent state.
  void IntegrateNewMeasurement(const imu::ImuCalibration& calib, const imu::ImuMeasurement& m) {
    measurements_.push_back(m);
    load(record_at(gyro_bias_, acc_bias_, calib.view()));
  }
  void Reintegrate(const imu::ImuCalibration& calib) {  // at the updated biases (imu_preintegration.cpp:129-143)
    load(record_at(gyro_bias_updated_, acc_bias_updated_, calib.view()));
  }
  void SetNewBias(const Vector3T& new_gyro_bias, const Vector3T& new_acc_bias) {
    gyro_bias_updated_ = new_gyro_bias;
    acc_bias_updated_ = new_acc_bias;
  }
  void GetDeltaBias(const Vector3T& new_gyro_bias, const Vector3T& new_acc_bias, Vector3T& gyro_bias_delta,
                    Vector3T& acc_bias_delta) const {
    for (int i = 0; i < 3; ++i) {
      gyro_bias_delta(i) = new_gyro_bias(i) - gyro_bias_(i);
      acc_bias_delta(i) = new_acc_bias(i) - acc_bias_(i);
    }
  }
  void GetDeltaBias(Vector3T& gyro_bias_delta, Vector3T& acc_bias_delta) const {
    GetDeltaBias(gyro_bias_updated_, acc_bias_updated_, gyro_bias_delta, acc_bias_delta);
  }
  // dR Exp(JRg (bg - bg0)), re-orthonormalized (imu_preintegration.cpp:162-166)
  Matrix3T GetDeltaRotation(const Vector3T& new_gyro_bias) const {
    double w[3], E[9], X[9];
    for (int r = 0; r < 3; ++r) {
      w[r] = 0.0;
      for (int c = 0; c < 3; ++c) w[r] += double(JRg(r, c)) * (new_gyro_bias(c) - gyro_bias_(c));
    }
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double a = th < 1e-12 ? 1.0 : std::sin(th) / th, b = th < 1e-12 ? 0.5 : (1.0 - std::cos(th)) / (th * th);
    const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double k2 = 0.0;
        for (int q = 0; q < 3; ++q) k2 += K[3 * r + q] * K[3 * q + c];
        E[3 * r + c] = (r == c ? 1.0 : 0.0) + a * K[3 * r + c] + b * k2;
      }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        X[3 * r + c] = 0.0;
        for (int q = 0; q < 3; ++q) X[3 * r + c] += double(dR(r, q)) * E[3 * q + c];
      }
    for (int it = 0; it < 6; ++it) {  // orthogonal polar factor, X <- (X + X^-T) / 2
      const double C[9] = {X[4] * X[8] - X[5] * X[7], X[5] * X[6] - X[3] * X[8], X[3] * X[7] - X[4] * X[6],
                           X[2] * X[7] - X[1] * X[8], X[0] * X[8] - X[2] * X[6], X[1] * X[6] - X[0] * X[7],
                           X[1] * X[5] - X[2] * X[4], X[2] * X[3] - X[0] * X[5], X[0] * X[4] - X[1] * X[3]};
      const double det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
      for (int i = 0; i < 9; ++i) X[i] = 0.5 * X[i] + 0.5 * C[i] / det;
    }
    Matrix3T R;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) R(r, c) = float(X[3 * r + c]);
    return R;
  }
  // dV + JVg (bg - bg0) + JVa (ba - ba0), dP likewise (imu_preintegration.cpp:168-181)
  Vector3T GetDeltaVelocity(const Vector3T& new_gyro_bias, const Vector3T& new_acc_bias) const {
    return corrected(dV, JVg, JVa, new_gyro_bias, new_acc_bias);
  }
  Vector3T GetDeltaPosition(const Vector3T& new_gyro_bias, const Vector3T& new_acc_bias) const {
    return corrected(dP, JPg, JPa, new_gyro_bias, new_acc_bias);
  }
  float GetDeltaT_s() const { return dT_s; }
  const Vector3T& GetOriginalGyroBias() const { return gyro_bias_; }
  const Vector3T& GetOriginalAccBias() const { return acc_bias_; }
  void GetUpdatedBias(Vector3T& gyro_bias_updated, Vector3T& acc_bias_updated) const {
    gyro_bias_updated = gyro_bias_updated_;
    acc_bias_updated = acc_bias_updated_;
  }
  void InfoMatrix(Matrix9T& info) const {
    Matrix3T g, a;
    information(info, g, a);
  }
  void InfoGyroRWMatrix(Matrix3T& info) const {
    Matrix9T i9;
    Matrix3T a;
    information(i9, info, a);
  }
  void InfoAccRWMatrix(Matrix3T& info) const {
    Matrix9T i9;
    Matrix3T g;
    information(i9, g, info);
  }
  bool empty() const { return measurements_.empty(); }
  size_t size() const { return measurements_.size(); }
  const std::vector<imu::ImuMeasurement>& measurements() const { return measurements_; }

  // The ABI record of this preintegration; its samples start at `first` of the problem's sample array.
  sba_imu_preintegration record(int32_t first) const {
    sba_imu_preintegration s{};
    std::memcpy(s.cov, cov_matrix_.m, sizeof(s.cov));
    std::memcpy(s.acc_rw_cov, acc_random_walk_accum_cov_matrix_.m, sizeof(s.acc_rw_cov));
    std::memcpy(s.gyro_rw_cov, gyro_random_walk_accum_cov_matrix_.m, sizeof(s.gyro_rw_cov));
    std::memcpy(s.dR, dR.m, sizeof(s.dR));
    std::memcpy(s.dV, dV.data(), sizeof(s.dV));
    std::memcpy(s.dP, dP.data(), sizeof(s.dP));
    std::memcpy(s.JRg, JRg.m, sizeof(s.JRg));
    std::memcpy(s.JVg, JVg.m, sizeof(s.JVg));
    std::memcpy(s.JVa, JVa.m, sizeof(s.JVa));
    std::memcpy(s.JPg, JPg.m, sizeof(s.JPg));
    std::memcpy(s.JPa, JPa.m, sizeof(s.JPa));
    s.dT_s = dT_s;
    std::memcpy(s.gyro_bias, gyro_bias_.data(), 12);
    std::memcpy(s.acc_bias, acc_bias_.data(), 12);
    std::memcpy(s.gyro_bias_updated, gyro_bias_updated_.data(), 12);
    std::memcpy(s.acc_bias_updated, acc_bias_updated_.data(), 12);
    s.first_measurement = first;
    s.num_measurements = static_cast<int32_t>(measurements_.size());
    return s;
  }
  // Takes the state of a record the solver returned (the samples stay).
  void load(const sba_imu_preintegration& s) {
    std::memcpy(cov_matrix_.m, s.cov, sizeof(s.cov));
    std::memcpy(acc_random_walk_accum_cov_matrix_.m, s.acc_rw_cov, sizeof(s.acc_rw_cov));
    std::memcpy(gyro_random_walk_accum_cov_matrix_.m, s.gyro_rw_cov, sizeof(s.gyro_rw_cov));
    std::memcpy(dR.m, s.dR, sizeof(s.dR));
    std::memcpy(dV.data(), s.dV, sizeof(s.dV));
    std::memcpy(dP.data(), s.dP, sizeof(s.dP));
    std::memcpy(JRg.m, s.JRg, sizeof(s.JRg));
    std::memcpy(JVg.m, s.JVg, sizeof(s.JVg));
    std::memcpy(JVa.m, s.JVa, sizeof(s.JVa));
    std::memcpy(JPg.m, s.JPg, sizeof(s.JPg));
    std::memcpy(JPa.m, s.JPa, sizeof(s.JPa));
    dT_s = s.dT_s;
    std::memcpy(gyro_bias_.data(), s.gyro_bias, 12);
    std::memcpy(acc_bias_.data(), s.acc_bias, 12);
    std::memcpy(gyro_bias_updated_.data(), s.gyro_bias_updated, 12);
    std::memcpy(acc_bias_updated_.data(), s.acc_bias_updated, 12);
  }

private:
  Vector3T corrected(const Vector3T& d, const Matrix3T& Jg, const Matrix3T& Ja, const Vector3T& bg,
                     const Vector3T& ba) const {
    Vector3T out;
    for (int r = 0; r < 3; ++r) {
      float s = d(r);
      for (int c = 0; c < 3; ++c) s += Jg(r, c) * (bg(c) - gyro_bias_(c)) + Ja(r, c) * (ba(c) - acc_bias_(c));
      out(r) = s;
    }
    return out;
  }

  std::vector<imu::ImuMeasurement> measurements_;
  Vector3T gyro_bias_, acc_bias_, gyro_bias_updated_, acc_bias_updated_;

  sba_imu_preintegration record_at(const Vector3T& bg, const Vector3T& ba, const sba_imu_calibration& c) const {
    std::vector<sba_imu_measurement> m(measurements_.size());
    for (size_t i = 0; i < m.size(); ++i) {
      m[i].time_ns = measurements_[i].time_ns;
      std::memcpy(m[i].linear_acceleration, measurements_[i].linear_acceleration.data(), 12);
      std::memcpy(m[i].angular_velocity, measurements_[i].angular_velocity.data(), 12);
    }
    sba_imu_preintegration s{};
    sba_imu_integrate(&c, m.data(), static_cast<int32_t>(m.size()), bg.data(), ba.data(), &s);
    return s;
  }
  void information(Matrix9T& info, Matrix3T& g, Matrix3T& a) const {
    const sba_imu_preintegration s = record(0);
    sba_imu_information(&s, info.m, g.m, a.m);
  }
};

// use IMU frame to simplify jacobians (imu_sba_problem.h:30-46)
struct Pose {
  Isometry3T w_from_imu = Isometry3T::Identity();
  Vector3T velocity;
  Vector3T gyro_bias;
  Vector3T acc_bias;
  // contains the IMU samples from this key frame until the next one
  IMUPreintegration preintegration;
  Matrix15T info = Matrix15T::Zero();
};

// The in/out problem (imu_sba_problem.h:48-91). Observations are ordered by key frame, then by point id; the
// first num_fixed_key_frames key frames do not move.
struct ImuBAProblem {
  std::vector<Vector3T> points = {};
  std::vector<Vector2T> observation_xys = {};
  std::vector<Matrix2T> observation_infos = {};
  std::vector<Matrix2T> info_matrix = {};
  std::vector<int32_t> point_ids = {};
  std::vector<int32_t> pose_ids = {};
  std::vector<int8_t> camera_ids = {};
  std::vector<Pose> rig_poses = {};
  int32_t num_fixed_key_frames = 0;
  camera::Rig rig = {};
  Vector3T gravity;  // world-frame gravity acceleration

  // solver options
  int32_t max_iterations = 7;
  float robustifier_scale = 1.f;
  float robustifier_scale_pose = 10.f;
  float prior_gyro = 0;
  float prior_acc = 0;
  float imu_penalty = 1;
  float acc_rw_penalty = -1;           // penalty for acc bias random walk; -1 means use imu_penalty
  float boundary_imu_penalty = 1e-2f;  // downweight IMU edge at fixed/optimizable boundary

  // solver statistic
  int32_t iterations = 0;
  float initial_cost = 0;
  float reintegration_thresh = 1e-4f;

  void Reset() {  // imu_sba_problem.cpp:49-65
    points.clear();
    observation_xys.clear();
    observation_infos.clear();
    info_matrix.clear();
    point_ids.clear();
    pose_ids.clear();
    camera_ids.clear();
    rig_poses.clear();
    num_fixed_key_frames = 0;
    gravity = Vector3T();
    max_iterations = 0;
    robustifier_scale = 0.5f;
    iterations = 0;
    initial_cost = 0;
    reintegration_thresh = 0;
  }
};

// IMUBundlerGpuFixedVel (libs/imu/imu_sba_gpu.h:42-47) on MI355X through sba_imu_solve: the problem is updated in
// place (points, key frame states, preintegrations, iterations, initial_cost); true when the final cost is below
// the initial one. No exception escapes; a device error returns false.
class IMUBundlerGpuFixedVel {
public:
  explicit IMUBundlerGpuFixedVel(const imu::ImuCalibration& calib, int device = 0);
  virtual ~IMUBundlerGpuFixedVel();
  IMUBundlerGpuFixedVel(const IMUBundlerGpuFixedVel&) = delete;
  IMUBundlerGpuFixedVel& operator=(const IMUBundlerGpuFixedVel&) = delete;
  bool solve(ImuBAProblem& problem);

private:
  imu::ImuCalibration calib_;
  class Impl;
  std::unique_ptr<Impl> impl;
};

}  // namespace cuvslam::sba_imu
