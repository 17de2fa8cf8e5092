// A caller of the reference's inertial-SBA API, compiled against the drop-in headers: reads a window written by
// tests/test_gpu_dropin_imu.py, builds cuvslam::sba_imu::ImuBAProblem with one IMUPreintegration per key frame fed
// sample by sample (IntegrateNewMeasurement, as the tracker does), calls IMUBundlerGpuFixedVel::solve and prints
// the solver statistics and the solved state as one JSON line.
//
// Input (little endian): int32 P N K F C; float points[P][3], xy[N][2], infos[N][4], int32 point_ids[N],
// pose_ids[N], camera_ids[N]; float camera_from_rig[C][16], w_from_imu[K][16], velocity[K][3], gyro_bias[K][3],
// acc_bias[K][3]; int32 samples[K]; per sample int64 time_ns, float acc[3], gyro[3]; float gravity[3],
// rig_from_imu[16], noise[5]; int32 max_iterations; float robustifier_scale, robustifier_scale_pose, prior_gyro,
// prior_acc, imu_penalty, acc_rw_penalty, boundary_imu_penalty, reintegration_thresh.
#include <cstdio>
#include <cstring>
#include <vector>

#include "sba/imu_sba.h"

using namespace cuvslam;

template <class T>
static bool get(FILE* f, T* dst, size_t n) {
  return std::fread(dst, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[5];
  if (!get(f, hdr, 5)) return 3;
  const int P = hdr[0], N = hdr[1], K = hdr[2], F = hdr[3], C = hdr[4];
  sba_imu::ImuBAProblem p;
  p.points.resize(P);
  p.observation_xys.resize(N);
  p.observation_infos.resize(N);
  p.point_ids.resize(N);
  p.pose_ids.resize(N);
  std::vector<int32_t> cid(N);
  bool ok = get(f, p.points[0].data(), 3 * size_t(P)) && get(f, p.observation_xys[0].data(), 2 * size_t(N)) &&
            get(f, p.observation_infos[0].data(), 4 * size_t(N)) && get(f, p.point_ids.data(), N) &&
            get(f, p.pose_ids.data(), N) && get(f, cid.data(), N);
  p.camera_ids.assign(cid.begin(), cid.end());
  p.rig.num_cameras = C;
  for (int c = 0; c < C && ok; ++c) ok = get(f, p.rig.camera_from_rig[c].data(), 16);
  p.rig_poses.resize(K);
  std::vector<float> buf(16 * size_t(K) + 9 * size_t(K));
  ok = ok && get(f, buf.data(), buf.size());
  std::vector<int32_t> ns(K);
  ok = ok && get(f, ns.data(), K);
  std::vector<std::vector<imu::ImuMeasurement>> samples(K);
  for (int k = 0; k < K && ok; ++k)
    for (int s = 0; s < ns[k] && ok; ++s) {
      imu::ImuMeasurement m;
      float v[6];
      ok = get(f, &m.time_ns, 1) && get(f, v, 6);
      for (int i = 0; i < 3; ++i) {
        m.linear_acceleration(i) = v[i];
        m.angular_velocity(i) = v[3 + i];
      }
      samples[k].push_back(m);
    }
  float g[3], rig_from_imu[16], noise[5];
  ok = ok && get(f, g, 3) && get(f, rig_from_imu, 16) && get(f, noise, 5);
  int32_t max_it = 0;
  float opt[8];
  ok = ok && get(f, &max_it, 1) && get(f, opt, 8);
  std::fclose(f);
  if (!ok) return 3;
  Isometry3T T;
  std::memcpy(T.data(), rig_from_imu, sizeof(rig_from_imu));
  const imu::ImuCalibration calib(T, noise[0], noise[1], noise[2], noise[3], noise[4]);
  for (int k = 0; k < K; ++k) {
    sba_imu::Pose& x = p.rig_poses[k];
    std::memcpy(x.w_from_imu.data(), &buf[16 * k], 64);
    for (int i = 0; i < 3; ++i) {
      x.velocity(i) = buf[16 * K + 3 * k + i];
      x.gyro_bias(i) = buf[19 * K + 3 * k + i];
      x.acc_bias(i) = buf[22 * K + 3 * k + i];
    }
    x.preintegration = sba_imu::IMUPreintegration(Vector3T(), Vector3T());
    for (const auto& m : samples[k]) x.preintegration.IntegrateNewMeasurement(calib, m);
  }
  p.num_fixed_key_frames = F;
  p.gravity = Vector3T(g[0], g[1], g[2]);
  p.max_iterations = max_it;
  p.robustifier_scale = opt[0];
  p.robustifier_scale_pose = opt[1];
  p.prior_gyro = opt[2];
  p.prior_acc = opt[3];
  p.imu_penalty = opt[4];
  p.acc_rw_penalty = opt[5];
  p.boundary_imu_penalty = opt[6];
  p.reintegration_thresh = opt[7];

  sba_imu::IMUBundlerGpuFixedVel bundler(calib);
  const bool solved = bundler.solve(p);
  std::printf("{\"ok\": %s, \"iterations\": %d, \"initial_cost\": %.9g, \"p0\": [%.9g, %.9g, %.9g], \"t\": [",
              solved ? "true" : "false", p.iterations, p.initial_cost, p.points[0](0), p.points[0](1),
              p.points[0](2));
  for (int k = 0; k < K; ++k)
    std::printf("%s[%.9g, %.9g, %.9g]", k ? ", " : "", p.rig_poses[k].w_from_imu(0, 3),
                p.rig_poses[k].w_from_imu(1, 3), p.rig_poses[k].w_from_imu(2, 3));
  std::printf("], \"bg\": [");
  for (int k = 0; k < K; ++k)
    std::printf("%s[%.9g, %.9g, %.9g]", k ? ", " : "", p.rig_poses[k].gyro_bias(0), p.rig_poses[k].gyro_bias(1),
                p.rig_poses[k].gyro_bias(2));
  Vector3T og = p.rig_poses[F].preintegration.GetOriginalGyroBias();
  std::printf("], \"pre_bg\": [%.9g, %.9g, %.9g], \"pre_samples\": %zu}\n", og(0), og(1), og(2),
              p.rig_poses[F].preintegration.size());
  return 0;
}
