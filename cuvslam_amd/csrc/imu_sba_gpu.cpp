// Drop-in cuvslam::sba_imu::IMUBundlerGpuFixedVel (libs/imu/imu_sba_gpu.h:42-47) over the C ABI of
// cuvslam_sba_imu.h: the problem's vectors are handed over as plain buffers (the key frame states and the
// preintegration records are gathered, the IMU samples of every key frame concatenated), and the solved state,
// preintegrations and statistics are written back in place.
#include <vector>

#include "cuvslam_sba.h"
#include "cuvslam_sba_imu.h"
#include "sba/imu_sba.h"

namespace cuvslam::sba_imu {

class IMUBundlerGpuFixedVel::Impl {
public:
  sba_handle* h = nullptr;
};

IMUBundlerGpuFixedVel::IMUBundlerGpuFixedVel(const imu::ImuCalibration& calib, int device)
    : calib_(calib), impl(new Impl) {
  sba_options o;
  sba_default_options(&o);
  o.device = device;
  if (sba_create(&o, &impl->h) != SBA_OK) impl->h = nullptr;
}

IMUBundlerGpuFixedVel::~IMUBundlerGpuFixedVel() {
  if (impl && impl->h) sba_destroy(impl->h);
}

bool IMUBundlerGpuFixedVel::solve(ImuBAProblem& problem) {
  if (!impl->h) return false;
  const int K = static_cast<int>(problem.rig_poses.size());
  std::vector<float> poses(16 * size_t(K)), vel(3 * size_t(K)), bg(3 * size_t(K)), ba(3 * size_t(K));
  std::vector<sba_imu_preintegration> pre(K);
  std::vector<sba_imu_measurement> meas;
  for (int k = 0; k < K; ++k) {
    const Pose& p = problem.rig_poses[k];
    for (int i = 0; i < 16; ++i) poses[16 * k + i] = p.w_from_imu.data()[i];
    for (int i = 0; i < 3; ++i) {
      vel[3 * k + i] = p.velocity(i);
      bg[3 * k + i] = p.gyro_bias(i);
      ba[3 * k + i] = p.acc_bias(i);
    }
    pre[k] = p.preintegration.record(static_cast<int32_t>(meas.size()));
    for (const auto& m : p.preintegration.measurements()) {
      sba_imu_measurement s{};
      s.time_ns = m.time_ns;
      for (int i = 0; i < 3; ++i) {
        s.linear_acceleration[i] = m.linear_acceleration(i);
        s.angular_velocity[i] = m.angular_velocity(i);
      }
      meas.push_back(s);
    }
  }
  sba_imu_problem_view v{};
  v.num_points = static_cast<int32_t>(problem.points.size());
  v.num_observations = static_cast<int32_t>(problem.observation_xys.size());
  v.num_poses = K;
  v.num_fixed_key_frames = problem.num_fixed_key_frames;
  v.num_cameras = problem.rig.num_cameras;
  v.num_measurements = static_cast<int32_t>(meas.size());
  v.points = problem.points.empty() ? nullptr : problem.points[0].data();
  v.observation_xys = problem.observation_xys.empty() ? nullptr : problem.observation_xys[0].data();
  v.observation_infos = problem.observation_infos.empty() ? nullptr : problem.observation_infos[0].data();
  v.point_ids = problem.point_ids.data();
  v.pose_ids = problem.pose_ids.data();
  v.camera_ids = problem.camera_ids.data();
  v.camera_from_rig = problem.rig.camera_from_rig[0].data();
  v.w_from_imu = poses.data();
  v.velocity = vel.data();
  v.gyro_bias = bg.data();
  v.acc_bias = ba.data();
  v.preintegration = pre.data();
  v.measurements = meas.data();
  for (int i = 0; i < 3; ++i) v.gravity[i] = problem.gravity(i);
  v.max_iterations = problem.max_iterations;
  v.robustifier_scale = problem.robustifier_scale;
  v.robustifier_scale_pose = problem.robustifier_scale_pose;
  v.prior_gyro = problem.prior_gyro;
  v.prior_acc = problem.prior_acc;
  v.imu_penalty = problem.imu_penalty;
  v.acc_rw_penalty = problem.acc_rw_penalty;
  v.boundary_imu_penalty = problem.boundary_imu_penalty;
  v.reintegration_thresh = problem.reintegration_thresh;

  std::vector<Vector3T> pts(problem.points.size());
  sba_imu_result r{};
  r.points = pts.empty() ? nullptr : pts[0].data();
  r.w_from_imu = poses.data();
  r.velocity = vel.data();
  r.gyro_bias = bg.data();
  r.acc_bias = ba.data();
  r.preintegration = pre.data();
  const sba_imu_calibration c = calib_.view();
  const int rc = sba_imu_solve(impl->h, &c, &v, &r);
  if (rc == SBA_E_DEVICE || rc == SBA_E_INVALID || rc == SBA_E_CAPACITY) return false;
  problem.iterations = r.iterations;
  problem.initial_cost = r.initial_cost;
  if (rc == SBA_OK || rc == SBA_E_NO_PROGRESS) {
    problem.points = pts;
    for (int k = 0; k < K; ++k) {
      Pose& p = problem.rig_poses[k];
      for (int i = 0; i < 16; ++i) p.w_from_imu.data()[i] = poses[16 * k + i];
      for (int i = 0; i < 3; ++i) {
        p.velocity(i) = vel[3 * k + i];
        p.gyro_bias(i) = bg[3 * k + i];
        p.acc_bias(i) = ba[3 * k + i];
      }
      p.preintegration.load(pre[k]);
    }
  }
  return rc == SBA_OK;
}

}  // namespace cuvslam::sba_imu
