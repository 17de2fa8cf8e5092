// Host runtime of the inertial SBA: the C ABI of include/cuvslam_sba_imu.h.
//
// sba_imu_solve follows IMUBundlerGpuFixedVel::solve (libs/imu/imu_sba_gpu_fixed_vel.cpp:75-363): the window is
// packed once, the kernels of sba_imu.hip run one LM iteration per host round trip (the host reads the iteration's
// scalars and decides, as the reference does), and after an accepted step the preintegrations of the free key frames
// are re-integrated on the host at their new biases when those moved past the threshold (imu_preintegration.cpp is
// host code in the reference too). The preintegration is computed in FP64 and stored in the ABI's FP32 record.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuvslam_sba_imu.h"
#include "sba_imu_device.h"
#include "sba_internal.h"

namespace sba_imu_dev {
hipError_t imu_model(const Window& w, hipStream_t s);
hipError_t imu_step(const Window& w, hipStream_t s);
hipError_t imu_cost(const Window& w, hipStream_t s);
hipError_t imu_accept(const Window& w, hipStream_t s);
}  // namespace sba_imu_dev

using namespace sba_imu_dev;

namespace {

constexpr float kFltEps = std::numeric_limits<float>::epsilon();  // common/types.h:203-206

// ---- 3x3 FP64 host helpers (row-major) ----------------------------------------------------------------------------

using M3 = std::array<double, 9>;
using V3 = std::array<double, 3>;

M3 mul(const M3& A, const M3& B) {
  M3 C{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
  return C;
}
V3 mul(const M3& A, const V3& x) {
  return V3{A[0] * x[0] + A[1] * x[1] + A[2] * x[2], A[3] * x[0] + A[4] * x[1] + A[5] * x[2],
            A[6] * x[0] + A[7] * x[1] + A[8] * x[2]};
}
M3 transpose(const M3& A) { return M3{A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]}; }
M3 skew(const V3& w) { return M3{0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}; }
M3 eye() { return M3{1, 0, 0, 0, 1, 0, 0, 0, 1}; }

M3 so3_exp(const V3& w) {  // Rodrigues
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const M3 S = skew(w), S2 = mul(S, S);
  M3 R = eye();
  if (th < 1e-12) {
    for (int i = 0; i < 9; ++i) R[i] += S[i];
    return R;
  }
  const double a = std::sin(th) / th, b = (1.0 - std::cos(th)) / (th * th);
  for (int i = 0; i < 9; ++i) R[i] += a * S[i] + b * S2[i];
  return R;
}

M3 so3_right_jacobian(const V3& x) {  // J_r(x) = J_l(-x) (math/twist.cpp:103-128)
  const V3 y{-x[0], -x[1], -x[2]};
  const double phi = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  const M3 S = skew(y);
  M3 J = eye();
  if (phi < 1e-8) {
    for (int i = 0; i < 9; ++i) J[i] += 0.5 * S[i];
    return J;
  }
  const double s = std::sin(phi) / phi, k = (1.0 - std::cos(phi)) / phi;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      J[3 * r + c] = (r == c ? s : 0.0) + (1.0 - s) * y[r] * y[c] / (phi * phi) + k * S[3 * r + c] / phi;
  return J;
}

// Orthogonal polar factor by Newton's iteration (U V^T of the SVD, common/rotation_utils.h:36-50).
M3 nearest_rotation(M3 X) {
  for (int it = 0; it < 8; ++it) {
    M3 C{X[4] * X[8] - X[5] * X[7], X[5] * X[6] - X[3] * X[8], X[3] * X[7] - X[4] * X[6],
         X[2] * X[7] - X[1] * X[8], X[0] * X[8] - X[2] * X[6], X[1] * X[6] - X[0] * X[7],
         X[1] * X[5] - X[2] * X[4], X[2] * X[3] - X[0] * X[5], X[0] * X[4] - X[1] * X[3]};
    const double det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
    double d = 0.0;
    for (int i = 0; i < 9; ++i) {
      const double v = 0.5 * X[i] + 0.5 * C[i] / det;
      d = std::max(d, std::fabs(v - X[i]));
      X[i] = v;
    }
    if (d < 1e-16) break;
  }
  return X;
}

// ---- IMU preintegration (imu_preintegration.cpp), FP64 ------------------------------------------------------------

struct Calib {
  M3 R_ri;
  V3 t_ri;
  double gyro_noise, acc_noise, gyro_rw, acc_rw, freq;  // isotropic covariances (imu_calibration.cpp:35-47)
};

Calib calib_of(const sba_imu_calibration& c) {
  Calib k{};
  for (int r = 0; r < 3; ++r) {
    for (int q = 0; q < 3; ++q) k.R_ri[3 * r + q] = c.rig_from_imu[4 * q + r];
    k.t_ri[r] = c.rig_from_imu[12 + r];
  }
  k.freq = c.frequency;
  k.gyro_noise = double(c.gyroscope_noise_density) * c.gyroscope_noise_density * k.freq;
  k.acc_noise = double(c.accelerometer_noise_density) * c.accelerometer_noise_density * k.freq;
  k.gyro_rw = double(c.gyroscope_random_walk) * c.gyroscope_random_walk / k.freq;
  k.acc_rw = double(c.accelerometer_random_walk) * c.accelerometer_random_walk / k.freq;
  return k;
}

struct Pre {
  double cov[81], cov_arw[9], cov_grw[9];
  M3 dR, JRg, JVg, JVa, JPg, JPa;
  V3 dV, dP;
  double dT;
  V3 bg, ba, bgu, bau;
  int first, n;
};

// Initialize + IntegrateNewMeasurement of samples[0..n) at biases (bg, ba) (imu_preintegration.cpp:40-127):
// the first step is 1/frequency, later ones the time difference; each update uses the deltas before it.
void integrate(Pre& p, const Calib& k, const sba_imu_measurement* m, int n, const V3& bg, const V3& ba) {
  std::memset(p.cov, 0, sizeof(p.cov));
  std::memset(p.cov_arw, 0, sizeof(p.cov_arw));
  std::memset(p.cov_grw, 0, sizeof(p.cov_grw));
  p.dR = eye();
  p.JRg = p.JVg = p.JVa = p.JPg = p.JPa = M3{};
  p.dV = p.dP = V3{};
  p.dT = 0.0;
  p.bg = p.bgu = bg;
  p.ba = p.bau = ba;
  for (int s = 0; s < n; ++s) {
    const double dt = s == 0 ? 1.0 / k.freq : double(m[s].time_ns - m[s - 1].time_ns) * 1e-9;
    const double dt2 = dt * dt;
    const V3 a{m[s].linear_acceleration[0] - ba[0], m[s].linear_acceleration[1] - ba[1],
               m[s].linear_acceleration[2] - ba[2]};
    const V3 w{m[s].angular_velocity[0] - bg[0], m[s].angular_velocity[1] - bg[1], m[s].angular_velocity[2] - bg[2]};
    const M3 dR = p.dR, dRsa = mul(dR, skew(a));
    const V3 Ra = mul(dR, a);
    for (int r = 0; r < 3; ++r) {
      p.dP[r] += p.dV[r] * dt + 0.5 * Ra[r] * dt2;
      p.dV[r] += Ra[r] * dt;
    }
    // state transition A (9x9) and noise maps B (acc), C (gyro) of [rotation, velocity, translation]
    double A[81] = {}, B[27] = {}, C[27] = {};
    for (int i = 0; i < 9; ++i) A[10 * i] = 1.0;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) {
        A[9 * (3 + r) + c] = -dt * dRsa[3 * r + c];
        A[9 * (6 + r) + c] = -0.5 * dt2 * dRsa[3 * r + c];
        B[3 * (3 + r) + c] = dt * dR[3 * r + c];
        B[3 * (6 + r) + c] = 0.5 * dt2 * dR[3 * r + c];
      }
      A[9 * (6 + r) + 3 + r] = dt;
    }
    const M3 dRsaJ = mul(dRsa, p.JRg);
    for (int i = 0; i < 9; ++i) {
      p.JPa[i] += p.JVa[i] * dt - 0.5 * dt2 * dR[i];
      p.JPg[i] += p.JVg[i] * dt - 0.5 * dt2 * dRsaJ[i];
      p.JVa[i] -= dt * dR[i];
      p.JVg[i] -= dt * dRsaJ[i];
    }
    const V3 wdt{w[0] * dt, w[1] * dt, w[2] * dt};
    const M3 E = so3_exp(wdt), Jr = so3_right_jacobian(wdt), ET = transpose(E);
    p.dR = nearest_rotation(mul(dR, E));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        A[9 * r + c] = ET[3 * r + c];
        C[3 * r + c] = dt * Jr[3 * r + c];
      }
    double T[81], N[81];
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) {
        double s2 = 0.0;
        for (int q = 0; q < 9; ++q) s2 += A[9 * i + q] * p.cov[9 * q + j];
        T[9 * i + j] = s2;
      }
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) {
        double s2 = 0.0, bb = 0.0, cc = 0.0;
        for (int q = 0; q < 9; ++q) s2 += T[9 * i + q] * A[9 * j + q];
        for (int q = 0; q < 3; ++q) {
          bb += B[3 * i + q] * B[3 * j + q];
          cc += C[3 * i + q] * C[3 * j + q];
        }
        N[9 * i + j] = s2 + k.acc_noise * bb + k.gyro_noise * cc;
      }
    std::memcpy(p.cov, N, sizeof(N));
    for (int i = 0; i < 3; ++i) {
      p.cov_arw[4 * i] += k.acc_rw;
      p.cov_grw[4 * i] += k.gyro_rw;
    }
    const M3 EJ = mul(ET, p.JRg);
    for (int i = 0; i < 9; ++i) p.JRg[i] = EJ[i] - dt * Jr[i];
    p.dT += dt;
  }
}

// Inverse of a symmetric positive definite n x n matrix by Cholesky; zero (and false) when it is not.
bool spd_inverse(int n, const double* A, double* X) {
  std::vector<double> L(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = A[n * j + j];
    for (int q = 0; q < j; ++q) d -= L[n * j + q] * L[n * j + q];
    if (!(d > 0.0)) {
      std::fill(X, X + n * n, 0.0);
      return false;
    }
    L[n * j + j] = std::sqrt(d);
    for (int i = j + 1; i < n; ++i) {
      double s = A[n * i + j];
      for (int q = 0; q < j; ++q) s -= L[n * i + q] * L[n * j + q];
      L[n * i + j] = s / L[n * j + j];
    }
  }
  for (int c = 0; c < n; ++c) {
    std::vector<double> y(n);
    for (int i = 0; i < n; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int q = 0; q < i; ++q) s -= L[n * i + q] * y[q];
      y[i] = s / L[n * i + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int q = i + 1; q < n; ++q) s -= L[n * q + i] * X[n * q + c];
      X[n * i + c] = s / L[n * i + i];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) X[n * i + j] = X[n * j + i] = 0.5 * (X[n * i + j] + X[n * j + i]);
  return true;
}

// InfoMatrix / InfoGyroRWMatrix / InfoAccRWMatrix (imu_preintegration.cpp:190-240): zero when empty, else the
// inverses (zero when a covariance is not positive definite).
void information(const Pre& p, double* info, double* igrw, double* iarw) {
  if (p.n == 0) {
    std::fill(info, info + 81, 0.0);
    std::fill(igrw, igrw + 9, 0.0);
    std::fill(iarw, iarw + 9, 0.0);
    return;
  }
  spd_inverse(9, p.cov, info);
  spd_inverse(3, p.cov_grw, igrw);
  spd_inverse(3, p.cov_arw, iarw);
}

void from_abi(Pre& p, const sba_imu_preintegration& s) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) p.cov[9 * i + j] = s.cov[9 * j + i];
  auto m3 = [](M3& m, const float* a) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[3 * r + c] = a[3 * c + r];
  };
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      p.cov_arw[3 * r + c] = s.acc_rw_cov[3 * c + r];
      p.cov_grw[3 * r + c] = s.gyro_rw_cov[3 * c + r];
    }
  m3(p.dR, s.dR);
  m3(p.JRg, s.JRg);
  m3(p.JVg, s.JVg);
  m3(p.JVa, s.JVa);
  m3(p.JPg, s.JPg);
  m3(p.JPa, s.JPa);
  for (int i = 0; i < 3; ++i) {
    p.dV[i] = s.dV[i];
    p.dP[i] = s.dP[i];
    p.bg[i] = s.gyro_bias[i];
    p.ba[i] = s.acc_bias[i];
    p.bgu[i] = s.gyro_bias_updated[i];
    p.bau[i] = s.acc_bias_updated[i];
  }
  p.dT = s.dT_s;
  p.first = s.first_measurement;
  p.n = s.num_measurements;
}

void to_abi(const Pre& p, sba_imu_preintegration& s) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) s.cov[9 * j + i] = float(p.cov[9 * i + j]);
  auto m3 = [](const M3& m, float* a) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a[3 * c + r] = float(m[3 * r + c]);
  };
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      s.acc_rw_cov[3 * c + r] = float(p.cov_arw[3 * r + c]);
      s.gyro_rw_cov[3 * c + r] = float(p.cov_grw[3 * r + c]);
    }
  m3(p.dR, s.dR);
  m3(p.JRg, s.JRg);
  m3(p.JVg, s.JVg);
  m3(p.JVa, s.JVa);
  m3(p.JPg, s.JPg);
  m3(p.JPa, s.JPa);
  for (int i = 0; i < 3; ++i) {
    s.dV[i] = float(p.dV[i]);
    s.dP[i] = float(p.dP[i]);
    s.gyro_bias[i] = float(p.bg[i]);
    s.acc_bias[i] = float(p.ba[i]);
    s.gyro_bias_updated[i] = float(p.bgu[i]);
    s.acc_bias_updated[i] = float(p.bau[i]);
  }
  s.dT_s = float(p.dT);
  s.first_measurement = p.first;
  s.num_measurements = p.n;
}

// The kernels' copy of a preintegration: the deltas, the updated-minus-integration gyroscope bias and the
// information matrices.
Preint device_record(const Pre& p) {
  Preint d{};
  std::copy(p.dR.begin(), p.dR.end(), d.dR);
  std::copy(p.JRg.begin(), p.JRg.end(), d.JRg);
  std::copy(p.JVg.begin(), p.JVg.end(), d.JVg);
  std::copy(p.JVa.begin(), p.JVa.end(), d.JVa);
  std::copy(p.JPg.begin(), p.JPg.end(), d.JPg);
  std::copy(p.JPa.begin(), p.JPa.end(), d.JPa);
  for (int i = 0; i < 3; ++i) {
    d.dV[i] = p.dV[i];
    d.dP[i] = p.dP[i];
    d.bg0[i] = p.bg[i];
    d.ba0[i] = p.ba[i];
    d.dbg[i] = p.bgu[i] - p.bg[i];
  }
  d.dT = p.dT;
  information(p, d.info, d.igrw, d.iarw);
  return d;
}

double norm3(const V3& a, const V3& b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return std::sqrt(x * x + y * y + z * z);
}

// ---- packing ------------------------------------------------------------------------------------------------------

struct Packed {
  int P, N, K, F, n, C, npairs, nblocks;
  std::vector<double> xy, info, cam, pts, pose;
  std::vector<int32_t> opid, okid, ocid, point_obs, pobs, point_pair, pair_obs, pair_pose, pair_point, pose_obs, kobs,
      pose_pair, kpair, blk_jk, blk_tri;
  std::vector<int2> tri;
};

int validate(sba_handle* h, const sba_imu_calibration* calib, const sba_imu_problem_view* v) {
  using sba_internal::set_error;
  if (!calib || !v) return set_error(h, SBA_E_INVALID, "null calibration or problem");
  if (v->num_points < 0 || v->num_observations < 0 || v->num_poses < 0 || v->num_cameras <= 0 ||
      v->num_measurements < 0 || v->num_fixed_key_frames < 0)
    return set_error(h, SBA_E_INVALID, "negative size");
  if ((v->num_points && !v->points) || (v->num_observations && (!v->observation_xys || !v->observation_infos ||
                                                                !v->point_ids || !v->pose_ids || !v->camera_ids)) ||
      (v->num_poses && (!v->w_from_imu || !v->velocity || !v->gyro_bias || !v->acc_bias || !v->preintegration)) ||
      !v->camera_from_rig || (v->num_measurements && !v->measurements))
    return set_error(h, SBA_E_INVALID, "null buffer");
  if (!(calib->frequency > 0.f)) return set_error(h, SBA_E_INVALID, "calibration frequency must be positive");
  for (int o = 0; o < v->num_observations; ++o)
    if (v->point_ids[o] < 0 || v->point_ids[o] >= v->num_points || v->pose_ids[o] < 0 ||
        v->pose_ids[o] >= v->num_poses || v->camera_ids[o] < 0 || v->camera_ids[o] >= v->num_cameras)
      return set_error(h, SBA_E_INVALID, "observation " + std::to_string(o) + " indexes out of range");
  for (int k = 0; k < v->num_poses; ++k) {
    const auto& p = v->preintegration[k];
    if (p.num_measurements < 0 || p.first_measurement < 0 ||
        int64_t(p.first_measurement) + p.num_measurements > v->num_measurements)
      return set_error(h, SBA_E_INVALID, "preintegration " + std::to_string(k) + " samples out of range");
  }
  const int n = v->num_poses - std::min(v->num_fixed_key_frames, v->num_poses);
  if (n > kMaxFree)
    return set_error(h, SBA_E_CAPACITY, "inertial window has " + std::to_string(n) + " free key frames; at most " +
                                            std::to_string(kMaxFree));
  return SBA_OK;
}

void pack(const sba_imu_calibration& calib, const sba_imu_problem_view& v, Packed& k) {
  const Calib cb = calib_of(calib);
  k.P = v.num_points;
  k.N = v.num_observations;
  k.K = v.num_poses;
  k.F = std::min(v.num_fixed_key_frames, v.num_poses);
  k.n = k.K - k.F;
  k.C = v.num_cameras;
  k.xy.assign(v.observation_xys, v.observation_xys + 2 * size_t(k.N));
  k.info.resize(4 * size_t(k.N));
  for (int o = 0; o < k.N; ++o) {  // column-major in, row-major out
    const float* L = v.observation_infos + 4 * o;
    k.info[4 * o] = L[0];
    k.info[4 * o + 1] = L[2];
    k.info[4 * o + 2] = L[1];
    k.info[4 * o + 3] = L[3];
  }
  k.opid.assign(v.point_ids, v.point_ids + k.N);
  k.okid.assign(v.pose_ids, v.pose_ids + k.N);
  k.ocid.resize(k.N);
  for (int o = 0; o < k.N; ++o) k.ocid[o] = v.camera_ids[o];
  k.cam.resize(12 * size_t(k.C));
  for (int c = 0; c < k.C; ++c) {  // cam_from_imu = camera_from_rig * rig_from_imu
    const float* T = v.camera_from_rig + 16 * c;
    M3 R;
    V3 t;
    for (int r = 0; r < 3; ++r) {
      for (int q = 0; q < 3; ++q) R[3 * r + q] = T[4 * q + r];
      t[r] = T[12 + r];
    }
    const M3 Rc = mul(R, cb.R_ri);
    const V3 tc = mul(R, cb.t_ri);
    for (int i = 0; i < 9; ++i) k.cam[12 * c + i] = Rc[i];
    for (int i = 0; i < 3; ++i) k.cam[12 * c + 9 + i] = tc[i] + t[i];
  }
  k.pts.assign(v.points, v.points + 3 * size_t(k.P));
  k.pose.assign(kPS * size_t(k.K), 0.0);
  for (int q = 0; q < k.K; ++q) {
    double* X = &k.pose[kPS * q];
    const float* T = v.w_from_imu + 16 * q;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) X[3 * r + c] = T[4 * c + r];
      X[kT + r] = T[12 + r];
      X[kV + r] = v.velocity[3 * q + r];
      X[kBg + r] = v.gyro_bias[3 * q + r];
      X[kBa + r] = v.acc_bias[3 * q + r];
    }
  }
  // observations by point, then key frame, then camera, then input order
  k.point_obs.assign(k.P + 1, 0);
  for (int o = 0; o < k.N; ++o) ++k.point_obs[k.opid[o] + 1];
  for (int i = 0; i < k.P; ++i) k.point_obs[i + 1] += k.point_obs[i];
  k.pobs.resize(k.N);
  {
    std::vector<int32_t> fill(k.point_obs.begin(), k.point_obs.end() - 1);
    for (int o = 0; o < k.N; ++o) k.pobs[fill[k.opid[o]]++] = o;
  }
  for (int i = 0; i < k.P; ++i)
    std::stable_sort(k.pobs.begin() + k.point_obs[i], k.pobs.begin() + k.point_obs[i + 1], [&](int a, int b) {
      return k.okid[a] != k.okid[b] ? k.okid[a] < k.okid[b] : k.ocid[a] < k.ocid[b];
    });
  // pairs: runs of one free key frame inside a point's observations
  k.point_pair.assign(k.P + 1, 0);
  k.pair_obs.clear();
  k.pair_pose.clear();
  k.pair_point.clear();
  for (int i = 0; i < k.P; ++i) {
    k.point_pair[i] = int32_t(k.pair_pose.size());
    for (int q = k.point_obs[i]; q < k.point_obs[i + 1];) {
      const int kf = k.okid[k.pobs[q]];
      int e = q + 1;
      while (e < k.point_obs[i + 1] && k.okid[k.pobs[e]] == kf) ++e;
      if (kf >= k.F) {
        k.pair_obs.push_back(q);
        k.pair_obs.push_back(e);
        k.pair_pose.push_back(kf - k.F);
        k.pair_point.push_back(i);
      }
      q = e;
    }
  }
  k.point_pair[k.P] = int32_t(k.pair_pose.size());
  k.npairs = int(k.pair_pose.size());
  // observations and pairs of each free key frame, in input and point order
  k.pose_obs.assign(k.n + 1, 0);
  for (int o = 0; o < k.N; ++o)
    if (k.okid[o] >= k.F) ++k.pose_obs[k.okid[o] - k.F + 1];
  for (int j = 0; j < k.n; ++j) k.pose_obs[j + 1] += k.pose_obs[j];
  k.kobs.assign(std::max(k.pose_obs[k.n], 1), 0);
  {
    std::vector<int32_t> fill(k.pose_obs.begin(), k.pose_obs.end() - 1);
    for (int o = 0; o < k.N; ++o)
      if (k.okid[o] >= k.F) k.kobs[fill[k.okid[o] - k.F]++] = o;
  }
  k.pose_pair.assign(k.n + 1, 0);
  for (int p = 0; p < k.npairs; ++p) ++k.pose_pair[k.pair_pose[p] + 1];
  for (int j = 0; j < k.n; ++j) k.pose_pair[j + 1] += k.pose_pair[j];
  k.kpair.assign(std::max(k.npairs, 1), 0);
  {
    std::vector<int32_t> fill(k.pose_pair.begin(), k.pose_pair.end() - 1);
    for (int p = 0; p < k.npairs; ++p) k.kpair[fill[k.pair_pose[p]]++] = p;
  }
  // lower blocks (j, k) of the reduced system and the pair products of each, in point order
  k.nblocks = k.n * (k.n + 1) / 2;
  k.blk_jk.resize(2 * size_t(std::max(k.nblocks, 1)));
  for (int j = 0; j < k.n; ++j)
    for (int q = 0; q <= j; ++q) {
      k.blk_jk[2 * (j * (j + 1) / 2 + q)] = j;
      k.blk_jk[2 * (j * (j + 1) / 2 + q) + 1] = q;
    }
  k.blk_tri.assign(k.nblocks + 1, 0);
  for (int i = 0; i < k.P; ++i)
    for (int a = k.point_pair[i]; a < k.point_pair[i + 1]; ++a)
      for (int b = k.point_pair[i]; b <= a; ++b) {
        const int j = k.pair_pose[a], q = k.pair_pose[b];
        ++k.blk_tri[j * (j + 1) / 2 + q + 1];
      }
  for (int b = 0; b < k.nblocks; ++b) k.blk_tri[b + 1] += k.blk_tri[b];
  k.tri.resize(std::max<size_t>(k.blk_tri[k.nblocks], 1));
  {
    std::vector<int32_t> fill(k.blk_tri.begin(), k.blk_tri.end() - 1);
    for (int i = 0; i < k.P; ++i)
      for (int a = k.point_pair[i]; a < k.point_pair[i + 1]; ++a)
        for (int b = k.point_pair[i]; b <= a; ++b) {
          const int j = k.pair_pose[a], q = k.pair_pose[b];
          k.tri[fill[j * (j + 1) / 2 + q]++] = int2{a, b};
        }
  }
}

// Device layout of one window in the handle's inertial workspace.
struct Layout {
  size_t off = 0;
  template <class T>
  size_t add(size_t count) {
    const size_t o = off;
    off += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
    return o;
  }
};

}  // namespace

extern "C" {

int sba_imu_integrate(const sba_imu_calibration* calib, const sba_imu_measurement* samples, int32_t n,
                      const float gyro_bias[3], const float acc_bias[3], sba_imu_preintegration* out) {
  if (!calib || !out || !gyro_bias || !acc_bias || n < 0 || (n > 0 && !samples) || !(calib->frequency > 0.f))
    return SBA_E_INVALID;
  Pre p{};
  integrate(p, calib_of(*calib), samples, n, V3{gyro_bias[0], gyro_bias[1], gyro_bias[2]},
            V3{acc_bias[0], acc_bias[1], acc_bias[2]});
  p.first = 0;
  p.n = n;
  to_abi(p, *out);
  return SBA_OK;
}

int sba_imu_information(const sba_imu_preintegration* pre, float info[81], float info_gyro_rw[9],
                        float info_acc_rw[9]) {
  if (!pre || !info || !info_gyro_rw || !info_acc_rw) return SBA_E_INVALID;
  Pre p{};
  from_abi(p, *pre);
  double i9[81], ig[9], ia[9];
  information(p, i9, ig, ia);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) info[9 * j + i] = float(i9[9 * i + j]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      info_gyro_rw[3 * j + i] = float(ig[3 * i + j]);
      info_acc_rw[3 * j + i] = float(ia[3 * i + j]);
    }
  return SBA_OK;
}

int sba_imu_solve(sba_handle* h, const sba_imu_calibration* calib, const sba_imu_problem_view* v,
                  sba_imu_result* res) {
  using sba_internal::set_error;
  if (!h) return SBA_E_INVALID;
  if (!res) return set_error(h, SBA_E_INVALID, "null result");
  if (int rc = validate(h, calib, v); rc != SBA_OK) return rc;
  const hipStream_t st = sba_internal::stream_of(h);
  auto hip_fail = [&](hipError_t e, const char* what) {
    return set_error(h, SBA_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  if (hipError_t e = hipSetDevice(sba_internal::device_of(h)); e != hipSuccess) return hip_fail(e, "hipSetDevice");

  Packed k;
  pack(*calib, *v, k);
  const Calib cb = calib_of(*calib);
  std::vector<Pre> pre(k.K);
  for (int q = 0; q < k.K; ++q) from_abi(pre[q], v->preintegration[q]);

  // outputs: the input state unless the solve moves it
  auto write_outputs = [&](const std::vector<double>& pts, const std::vector<double>& pose) {
    if (res->points)
      for (size_t i = 0; i < pts.size(); ++i) res->points[i] = float(pts[i]);
    for (int q = 0; q < k.K; ++q) {
      const double* X = &pose[kPS * q];
      if (res->w_from_imu) {
        float* T = res->w_from_imu + 16 * q;
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) T[4 * c + r] = float(X[3 * r + c]);
          T[12 + r] = float(X[kT + r]);
          T[4 * r + 3] = 0.f;
        }
        T[15] = 1.f;
      }
      for (int r = 0; r < 3; ++r) {
        if (res->velocity) res->velocity[3 * q + r] = float(X[kV + r]);
        if (res->gyro_bias) res->gyro_bias[3 * q + r] = float(X[kBg + r]);
        if (res->acc_bias) res->acc_bias[3 * q + r] = float(X[kBa + r]);
      }
      if (res->preintegration) to_abi(pre[q], res->preintegration[q]);
    }
  };
  res->iterations = 0;
  res->accepted = 0;
  res->lambda = 0.f;
  res->initial_cost = 0.f;
  res->initial_cost_f64 = res->last_cost_f64 = 0.0;
  // imu_sba_gpu_fixed_vel.cpp:78-82: no inertial edge, points or observations
  if (k.K - 1 <= 0 || k.P == 0 || k.N == 0) {
    write_outputs(k.pts, k.pose);
    res->status = SBA_E_EMPTY;
    return SBA_E_EMPTY;
  }

  // device block
  Layout L;
  const size_t n15 = 15 * size_t(k.n);
  const size_t o_xy = L.add<double>(2 * size_t(k.N)), o_info = L.add<double>(4 * size_t(k.N));
  const size_t o_opid = L.add<int32_t>(k.N), o_okid = L.add<int32_t>(k.N), o_ocid = L.add<int32_t>(k.N);
  const size_t o_cam = L.add<double>(k.cam.size());
  const size_t o_point_obs = L.add<int32_t>(k.point_obs.size()), o_pobs = L.add<int32_t>(k.pobs.size());
  const size_t o_point_pair = L.add<int32_t>(k.point_pair.size()), o_pair_obs = L.add<int32_t>(k.pair_obs.size());
  const size_t o_pair_pose = L.add<int32_t>(k.pair_pose.size()), o_pair_point = L.add<int32_t>(k.pair_point.size());
  const size_t o_pose_obs = L.add<int32_t>(k.pose_obs.size()), o_kobs = L.add<int32_t>(k.kobs.size());
  const size_t o_pose_pair = L.add<int32_t>(k.pose_pair.size()), o_kpair = L.add<int32_t>(k.kpair.size());
  const size_t o_blk_jk = L.add<int32_t>(k.blk_jk.size()), o_blk_tri = L.add<int32_t>(k.blk_tri.size());
  const size_t o_tri = L.add<int2>(k.tri.size());
  const size_t o_pre = L.add<Preint>(k.K);
  const size_t o_pts0 = L.add<double>(3 * size_t(k.P)), o_pts1 = L.add<double>(3 * size_t(k.P));
  const size_t o_pose0 = L.add<double>(kPS * size_t(k.K)), o_pose1 = L.add<double>(kPS * size_t(k.K));
  const size_t o_om = L.add<ObsModel>(k.N), o_em = L.add<EdgeModel>(k.K);
  const size_t o_U = L.add<double>(n15 * n15), o_ea = L.add<double>(n15);
  const size_t o_V = L.add<double>(9 * size_t(k.P)), o_eb = L.add<double>(3 * size_t(k.P));
  const size_t o_W = L.add<double>(18 * size_t(k.npairs)), o_Vinv = L.add<double>(9 * size_t(k.P));
  const size_t o_rp = L.add<double>(3 * size_t(k.P)), o_Y = L.add<double>(18 * size_t(k.npairs));
  const size_t o_A = L.add<double>(256 * size_t(std::max(k.nblocks, 1)));  // lower 16 x 16 tiles
  const size_t o_rhs = L.add<double>(n15), o_dp = L.add<double>(3 * size_t(k.P));
  const size_t o_ocost = L.add<double>(k.N), o_ecost = L.add<double>(k.K), o_ppred = L.add<double>(k.P);
  const size_t o_pmax = L.add<double>(k.P), o_cpred = L.add<double>(n15);
  const size_t o_out = L.add<IterScalars>(1), o_chol = L.add<int32_t>(1);
  char* base = nullptr;
  if (int rc = sba_internal::aux_workspace(h, L.off, &base); rc != SBA_OK) return rc;

  Window w{};
  w.P = k.P; w.N = k.N; w.K = k.K; w.F = k.F; w.n = k.n; w.C = k.C;
  w.npairs = k.npairs;
  w.nblocks = k.nblocks;
  for (int a = 0; a < 3; ++a) w.g[a] = v->gravity[a];
  w.scale = v->robustifier_scale;
  w.prior_gyro = v->prior_gyro;
  w.prior_acc = v->prior_acc;
  w.imu_penalty = v->imu_penalty;
  w.acc_rw_penalty = v->acc_rw_penalty;
  w.boundary_penalty = v->boundary_imu_penalty;
  auto at = [&](size_t o) { return base + o; };
  w.xy = reinterpret_cast<double*>(at(o_xy));
  w.info = reinterpret_cast<double*>(at(o_info));
  w.opid = reinterpret_cast<int32_t*>(at(o_opid));
  w.okid = reinterpret_cast<int32_t*>(at(o_okid));
  w.ocid = reinterpret_cast<int32_t*>(at(o_ocid));
  w.cam = reinterpret_cast<double*>(at(o_cam));
  w.point_obs = reinterpret_cast<int32_t*>(at(o_point_obs));
  w.pobs = reinterpret_cast<int32_t*>(at(o_pobs));
  w.point_pair = reinterpret_cast<int32_t*>(at(o_point_pair));
  w.pair_obs = reinterpret_cast<int32_t*>(at(o_pair_obs));
  w.pair_pose = reinterpret_cast<int32_t*>(at(o_pair_pose));
  w.pair_point = reinterpret_cast<int32_t*>(at(o_pair_point));
  w.pose_obs = reinterpret_cast<int32_t*>(at(o_pose_obs));
  w.kobs = reinterpret_cast<int32_t*>(at(o_kobs));
  w.pose_pair = reinterpret_cast<int32_t*>(at(o_pose_pair));
  w.kpair = reinterpret_cast<int32_t*>(at(o_kpair));
  w.blk_jk = reinterpret_cast<int32_t*>(at(o_blk_jk));
  w.blk_tri = reinterpret_cast<int32_t*>(at(o_blk_tri));
  w.tri = reinterpret_cast<int2*>(at(o_tri));
  Preint* dpre = reinterpret_cast<Preint*>(at(o_pre));
  w.pre = dpre;
  w.pts[0] = reinterpret_cast<double*>(at(o_pts0));
  w.pts[1] = reinterpret_cast<double*>(at(o_pts1));
  w.pose[0] = reinterpret_cast<double*>(at(o_pose0));
  w.pose[1] = reinterpret_cast<double*>(at(o_pose1));
  w.om = reinterpret_cast<ObsModel*>(at(o_om));
  w.em = reinterpret_cast<EdgeModel*>(at(o_em));
  w.U = reinterpret_cast<double*>(at(o_U));
  w.ea = reinterpret_cast<double*>(at(o_ea));
  w.V = reinterpret_cast<double*>(at(o_V));
  w.eb = reinterpret_cast<double*>(at(o_eb));
  w.W = reinterpret_cast<double*>(at(o_W));
  w.Vinv = reinterpret_cast<double*>(at(o_Vinv));
  w.rp = reinterpret_cast<double*>(at(o_rp));
  w.Y = reinterpret_cast<double*>(at(o_Y));
  w.A = reinterpret_cast<double*>(at(o_A));
  w.rhs = reinterpret_cast<double*>(at(o_rhs));
  w.dp = reinterpret_cast<double*>(at(o_dp));
  w.ocost = reinterpret_cast<double*>(at(o_ocost));
  w.ecost = reinterpret_cast<double*>(at(o_ecost));
  w.ppred = reinterpret_cast<double*>(at(o_ppred));
  w.pmax = reinterpret_cast<double*>(at(o_pmax));
  w.cpred = reinterpret_cast<double*>(at(o_cpred));
  w.out = reinterpret_cast<IterScalars*>(at(o_out));
  w.chol_ok = reinterpret_cast<int32_t*>(at(o_chol));

  auto up = [&](size_t o, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(at(o), src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
  };
  std::vector<Preint> hpre(k.K);
  for (int q = 0; q < k.K; ++q) hpre[q] = device_record(pre[q]);
  hipError_t e = hipSuccess;
  const std::pair<size_t, std::pair<const void*, size_t>> uploads[] = {
      {o_xy, {k.xy.data(), k.xy.size() * 8}},
      {o_info, {k.info.data(), k.info.size() * 8}},
      {o_opid, {k.opid.data(), k.opid.size() * 4}},
      {o_okid, {k.okid.data(), k.okid.size() * 4}},
      {o_ocid, {k.ocid.data(), k.ocid.size() * 4}},
      {o_cam, {k.cam.data(), k.cam.size() * 8}},
      {o_point_obs, {k.point_obs.data(), k.point_obs.size() * 4}},
      {o_pobs, {k.pobs.data(), k.pobs.size() * 4}},
      {o_point_pair, {k.point_pair.data(), k.point_pair.size() * 4}},
      {o_pair_obs, {k.pair_obs.data(), k.pair_obs.size() * 4}},
      {o_pair_pose, {k.pair_pose.data(), k.pair_pose.size() * 4}},
      {o_pair_point, {k.pair_point.data(), k.pair_point.size() * 4}},
      {o_pose_obs, {k.pose_obs.data(), k.pose_obs.size() * 4}},
      {o_kobs, {k.kobs.data(), k.kobs.size() * 4}},
      {o_pose_pair, {k.pose_pair.data(), k.pose_pair.size() * 4}},
      {o_kpair, {k.kpair.data(), k.kpair.size() * 4}},
      {o_blk_jk, {k.blk_jk.data(), k.blk_jk.size() * 4}},
      {o_blk_tri, {k.blk_tri.data(), k.blk_tri.size() * 4}},
      {o_tri, {k.tri.data(), k.tri.size() * sizeof(int2)}},
      {o_pre, {hpre.data(), hpre.size() * sizeof(Preint)}},
      {o_pts0, {k.pts.data(), k.pts.size() * 8}},
      {o_pts1, {k.pts.data(), k.pts.size() * 8}},
      {o_pose0, {k.pose.data(), k.pose.size() * 8}},
      {o_pose1, {k.pose.data(), k.pose.size() * 8}},
  };
  for (const auto& u : uploads)
    if (e == hipSuccess) e = up(u.first, u.second.first, u.second.second);
  if (e != hipSuccess) return hip_fail(e, "upload");

  IterScalars sc{};
  auto read_scalars = [&]() {
    hipError_t r = hipMemcpyAsync(&sc, w.out, sizeof(sc), hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    return r;
  };
  std::vector<double> cur_pose(k.pose), cur_pts(k.pts);

  // initial model and cost (imu_sba_gpu_fixed_vel.cpp:95-122)
  if ((e = imu_model(w, st)) != hipSuccess || (e = imu_cost(w, st)) != hipSuccess || (e = read_scalars()) != hipSuccess)
    return hip_fail(e, "initial cost");
  const double inf = std::numeric_limits<double>::infinity();
  const double initial = sc.skipped == k.N ? inf : sc.cost;
  res->initial_cost = float(initial);
  res->initial_cost_f64 = initial;
  res->last_cost_f64 = initial;
  if (initial < kFltEps) {
    write_outputs(k.pts, k.pose);
    res->status = SBA_OK;
    return SBA_OK;
  }
  if (!std::isfinite(initial)) {
    write_outputs(k.pts, k.pose);
    res->status = SBA_E_INFEASIBLE;
    return SBA_E_INFEASIBLE;
  }

  float lambda = 1e2f;
  double current = initial;
  int it = 0, accepted = 0;
  // SetNewBias of every free key frame and Reintegrate where its biases moved past the threshold
  // (imu_sba_gpu_fixed_vel.cpp:286-301); `from_original`: the final check after the loop (:339-356), which
  // re-integrates (and only then sets the new biases) where the biases moved from the integration's.
  auto reintegrate = [&](bool from_original) {
    // key frames are independent: spread over host threads (each writes only its own preintegration)
    const int nthreads = static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    sba_internal::parallel_for(h, k.K - k.F, nthreads, [&](int qq) {
      const int q = k.F + qq;
      const double* X = &cur_pose[kPS * q];
      const V3 bg{X[kBg], X[kBg + 1], X[kBg + 2]}, ba{X[kBa], X[kBa + 1], X[kBa + 2]};
      Pre& p = pre[q];
      const bool moved = norm3(bg, p.bg) > v->reintegration_thresh || norm3(ba, p.ba) > v->reintegration_thresh;
      if (!from_original || moved) {
        p.bgu = bg;
        p.bau = ba;
      }
      if (moved) {
        const int first = p.first, n = p.n;
        integrate(p, cb, v->measurements + first, n, bg, ba);
        p.first = first;
        p.n = n;
      }
    });
  };
  while (it < v->max_iterations) {
    ++it;
    w.lambda = lambda;
    if ((e = imu_step(w, st)) != hipSuccess || (e = read_scalars()) != hipSuccess) return hip_fail(e, "LM step");
    if (!sc.chol_ok) {  // the damped reduced system is not positive definite (:199-203)
      lambda *= 5.f;
      continue;
    }
    const double cost = sc.skipped == k.N ? inf : sc.cost;
    const double pred = sc.pred / current;
    if (!std::isfinite(cost)) {  // a non-finite trial cost is a bad step (:208-212)
      lambda *= 5.f;
      continue;
    }
    if (current < initial * kFltEps || (pred < kFltEps && !(sc.maxdp >= kFltEps) && !(sc.maxdc >= kFltEps))) break;
    const double rho = (1.0 - cost / current) / pred;
    if (rho > 0.25) {
      if (rho > 0.75 && lambda * 0.125f > 0.f) lambda *= 0.125f;
      current = cost;
      ++accepted;
      if ((e = imu_accept(w, st)) != hipSuccess) return hip_fail(e, "accept");
      e = hipMemcpyAsync(cur_pose.data(), w.pose[0], cur_pose.size() * 8, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(e, "download key frames");
      if (it == v->max_iterations) break;
      reintegrate(false);
      sba_internal::parallel_for(h, k.K - k.F, 16, [&](int qq) { hpre[k.F + qq] = device_record(pre[k.F + qq]); });
      if (k.n > 0) {
        e = hipMemcpyAsync(dpre + k.F, hpre.data() + k.F, size_t(k.n) * sizeof(Preint), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return hip_fail(e, "upload preintegration");
      }
      if ((e = imu_model(w, st)) != hipSuccess) return hip_fail(e, "model");
    } else {
      lambda *= 5.f;
    }
  }
  if (current < initial) reintegrate(true);
  e = hipMemcpyAsync(cur_pts.data(), w.pts[0], cur_pts.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(cur_pose.data(), w.pose[0], cur_pose.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "download state");
  write_outputs(cur_pts, cur_pose);
  res->iterations = it;
  res->accepted = accepted;
  res->lambda = lambda;
  res->last_cost_f64 = current;
  res->status = current < initial ? SBA_OK : SBA_E_NO_PROGRESS;
  return res->status;
}

}  // extern "C"
