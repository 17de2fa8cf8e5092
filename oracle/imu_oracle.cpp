// TEST INFRASTRUCTURE -- the CPU oracle of the inertial sliding-window SBA. Only tests/ and the smoke check load
// this library, and only as the checker; the product (cuvslam_amd/) never calls it.
//
// A restatement in plain FP64 C++ (no Eigen) of cuVSLAM's visual-inertial bundler with the GPU bundler's control
// flow and penalties:
//   * IMU preintegration           libs/imu/imu_preintegration.cpp:26-240
//   * model (residuals, Jacobians) libs/imu/imu_sba_fixed_vel.cpp:455-652 (update_model of sba_imu_v1.cu)
//   * full system                  imu_sba_fixed_vel.cpp:654-1004, with the penalties of the GPU assembly
//                                  (sba_imu_v1.cu:1521-1805: the accelerometer random walk is not down-weighted at
//                                  the fixed/free boundary)
//   * reduced system               imu_sba_fixed_vel.cpp:1006-1071 (point blocks inverted by LDLT as
//                                  sba_imu_v1.cu:247-294), Cholesky of the reduced pose system (cusolverDnSpotrf,
//                                  imu_sba_gpu_fixed_vel.cpp:383-397)
//   * cost, prediction             sba_imu_v1.cu:416-852
//   * LM loop with re-integration  imu_sba_gpu_fixed_vel.cpp:75-363
// Matrices are row-major here; the ABI structs (include/cuvslam_sba_imu.h) are column-major like Eigen, and are
// transposed on the way in and out. What pins each function is in tests/test_imu_oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/cuvslam_sba_imu.h"

namespace {

constexpr float kFltEps = std::numeric_limits<float>::epsilon();  // common/types.h:203-206

struct M3 {
  double a[9];
  double& operator()(int r, int c) { return a[3 * r + c]; }
  double operator()(int r, int c) const { return a[3 * r + c]; }
};
struct V3 {
  double v[3];
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};

M3 eye3() { M3 m{}; m(0, 0) = m(1, 1) = m(2, 2) = 1.0; return m; }
M3 mul(const M3& A, const M3& B) {
  M3 C{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
  return C;
}
V3 mul(const M3& A, const V3& x) {
  V3 y{};
  for (int r = 0; r < 3; ++r) y[r] = A(r, 0) * x[0] + A(r, 1) * x[1] + A(r, 2) * x[2];
  return y;
}
M3 tr(const M3& A) { M3 B{}; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B(r, c) = A(c, r); return B; }
M3 add(const M3& A, const M3& B, double s = 1.0) { M3 C{}; for (int i = 0; i < 9; ++i) C.a[i] = A.a[i] + s * B.a[i]; return C; }
M3 scale(const M3& A, double s) { M3 C{}; for (int i = 0; i < 9; ++i) C.a[i] = s * A.a[i]; return C; }
V3 addv(const V3& a, const V3& b, double s = 1.0) { return V3{{a[0] + s * b[0], a[1] + s * b[1], a[2] + s * b[2]}}; }
V3 scalev(const V3& a, double s) { return V3{{s * a[0], s * a[1], s * a[2]}}; }
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
M3 skew(const V3& w) {
  M3 m{};
  m(0, 1) = -w[2]; m(0, 2) = w[1];
  m(1, 0) = w[2];  m(1, 2) = -w[0];
  m(2, 0) = -w[1]; m(2, 1) = w[0];
  return m;
}
M3 outer(const V3& a, const V3& b) { M3 m{}; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) m(r, c) = a[r] * b[c]; return m; }

// SO(3) exponential (math/twist.cpp:75-82, Rodrigues).
M3 Exp(const V3& w) {
  const double th = norm(w);
  if (th < 1e-12) return add(eye3(), skew(w));
  const V3 n = scalev(w, 1.0 / th);
  const double c = std::cos(th), s = std::sin(th);
  return add(add(scale(eye3(), c), skew(scalev(n, s))), outer(n, n), 1.0 - c);
}

// SO(3) logarithm (math/twist.cpp:84-93; the GPU's Log, sba_imu_v1.cu:96-111): angle atan2(|2 sin|, 2 cos); near pi
// the axis comes from the symmetric part, where the skew part vanishes.
V3 Log(const M3& R) {
  V3 r{{R(2, 1) - R(1, 2), R(0, 2) - R(2, 0), R(1, 0) - R(0, 1)}};
  const double two_sin = norm(r), two_cos = R(0, 0) + R(1, 1) + R(2, 2) - 1.0;
  const double th = std::atan2(two_sin, two_cos);
  if (two_cos > -1.9) {
    const double s = th < 1e-6 ? 0.5 * (1.0 + th * th / 6.0) : th / two_sin;
    return scalev(r, s);
  }
  // R + R^T = 2 cos I + 2 (1 - cos) n n^T: the axis is the largest column of the symmetric part, signed by r
  M3 S = add(R, tr(R));
  for (int i = 0; i < 3; ++i) S(i, i) -= two_cos;
  int k = 0;
  for (int i = 1; i < 3; ++i) if (S(i, i) > S(k, k)) k = i;
  V3 n{{S(0, k), S(1, k), S(2, k)}};
  n = scalev(n, 1.0 / norm(n));
  if (dot(n, r) < 0) n = scalev(n, -1.0);
  return scalev(n, th);
}

// Left Jacobian of SO(3) and its inverse (math/twist.cpp:103-126); right = left of -x (:128-130).
M3 JlSO3(const V3& x) {
  const double phi = norm(x);
  if (phi < 1e-8) return add(eye3(), skew(x), 0.5);
  const V3 a = scalev(x, 1.0 / phi);
  const double s = std::sin(phi) / phi, k = (1.0 - std::cos(phi)) / phi;
  return add(add(scale(eye3(), s), outer(a, a), 1.0 - s), skew(a), k);
}
M3 JlInvSO3(const V3& x) {
  const double phi = norm(x);
  if (phi < 1e-8) return add(eye3(), skew(x), -0.5);
  const V3 a = scalev(x, 1.0 / phi);
  const double h = 0.5 * phi, c = h / std::tan(h);
  return add(add(scale(eye3(), c), outer(a, a), 1.0 - c), skew(a), -h);
}
M3 JrSO3(const V3& x) { return JlSO3(scalev(x, -1.0)); }
M3 JrInvSO3(const V3& x) { return JlInvSO3(scalev(x, -1.0)); }

// Nearest rotation (common/rotation_utils.h:36-50, U V^T of the SVD): the orthogonal polar factor by Newton's
// iteration X <- (X + X^-T) / 2, for a matrix with positive determinant.
M3 nearest_rotation(const M3& A) {
  M3 X = A;
  for (int it = 0; it < 30; ++it) {
    const double det = X(0, 0) * (X(1, 1) * X(2, 2) - X(1, 2) * X(2, 1)) -
                       X(0, 1) * (X(1, 0) * X(2, 2) - X(1, 2) * X(2, 0)) +
                       X(0, 2) * (X(1, 0) * X(2, 1) - X(1, 1) * X(2, 0));
    M3 C{};  // cofactor matrix = det * X^-T
    C(0, 0) = X(1, 1) * X(2, 2) - X(1, 2) * X(2, 1);
    C(0, 1) = X(1, 2) * X(2, 0) - X(1, 0) * X(2, 2);
    C(0, 2) = X(1, 0) * X(2, 1) - X(1, 1) * X(2, 0);
    C(1, 0) = X(0, 2) * X(2, 1) - X(0, 1) * X(2, 2);
    C(1, 1) = X(0, 0) * X(2, 2) - X(0, 2) * X(2, 0);
    C(1, 2) = X(0, 1) * X(2, 0) - X(0, 0) * X(2, 1);
    C(2, 0) = X(0, 1) * X(1, 2) - X(0, 2) * X(1, 1);
    C(2, 1) = X(0, 2) * X(1, 0) - X(0, 0) * X(1, 2);
    C(2, 2) = X(0, 0) * X(1, 1) - X(0, 1) * X(1, 0);
    M3 N = scale(add(X, scale(C, 1.0 / det)), 0.5);
    double d = 0;
    for (int i = 0; i < 9; ++i) d = std::max(d, std::fabs(N.a[i] - X.a[i]));
    X = N;
    if (d < 1e-16) break;
  }
  return X;
}

// ---- IMU preintegration (imu_preintegration.cpp) -----------------------------------------------------------------

struct Calib {
  M3 R_ri;  // rig_from_imu
  V3 t_ri;
  double gyro_noise, acc_noise, gyro_rw, acc_rw;  // isotropic covariances (imu_calibration.cpp:35-47)
  double freq;
};

Calib calib_of(const sba_imu_calibration& c) {
  Calib k{};
  for (int r = 0; r < 3; ++r) {
    for (int cc = 0; cc < 3; ++cc) k.R_ri(r, cc) = c.rig_from_imu[4 * cc + r];
    k.t_ri[r] = c.rig_from_imu[12 + r];
  }
  const double f = c.frequency;
  k.gyro_noise = double(c.gyroscope_noise_density) * c.gyroscope_noise_density * f;
  k.acc_noise = double(c.accelerometer_noise_density) * c.accelerometer_noise_density * f;
  k.gyro_rw = double(c.gyroscope_random_walk) * c.gyroscope_random_walk / f;
  k.acc_rw = double(c.accelerometer_random_walk) * c.accelerometer_random_walk / f;
  k.freq = f;
  return k;
}

struct Preint {
  double cov[81], cov_arw[9], cov_grw[9];
  M3 dR, JRg, JVg, JVa, JPg, JPa;
  V3 dV, dP;
  double dT;
  V3 bg, ba;    // biases of the integration
  V3 bgu, bau;  // updated biases (SetNewBias)
  int first, n;
};

// Initialize (imu_preintegration.cpp:40-67).
void preint_init(Preint& p, const V3& bg, const V3& ba) {
  std::memset(p.cov, 0, sizeof(p.cov));
  std::memset(p.cov_arw, 0, sizeof(p.cov_arw));
  std::memset(p.cov_grw, 0, sizeof(p.cov_grw));
  p.dR = eye3();
  p.JRg = p.JVg = p.JVa = p.JPg = p.JPa = M3{};
  p.dV = p.dP = V3{};
  p.dT = 0;
  p.bg = p.bgu = bg;
  p.ba = p.bau = ba;
}

// update_state (imu_preintegration.cpp:69-115): every term uses dR, JRg, JVa, JVg before this sample's update.
void preint_update(Preint& p, const Calib& k, const V3& acc, const V3& gyro, double dt) {
  const V3 a = addv(acc, p.ba, -1.0), w = addv(gyro, p.bg, -1.0);
  const double dt2 = dt * dt;
  const M3 dR = p.dR, sa = skew(a);
  double A[81] = {}, B[27] = {}, C[27] = {};  // A 9x9, B and C 9x3, row-major
  for (int i = 0; i < 9; ++i) A[10 * i] = 1.0;
  p.dP = addv(addv(p.dP, p.dV, dt), mul(dR, a), 0.5 * dt2);
  p.dV = addv(p.dV, mul(dR, a), dt);
  const M3 dRsa = mul(dR, sa);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[9 * (3 + r) + c] = -dt * dRsa(r, c);
      A[9 * (6 + r) + c] = -0.5 * dt2 * dRsa(r, c);
      B[3 * (3 + r) + c] = dt * dR(r, c);
      B[3 * (6 + r) + c] = 0.5 * dt2 * dR(r, c);
    }
  for (int r = 0; r < 3; ++r) A[9 * (6 + r) + 3 + r] = dt;
  p.JPa = add(add(p.JPa, p.JVa, dt), dR, -0.5 * dt2);
  p.JPg = add(add(p.JPg, p.JVg, dt), mul(dRsa, p.JRg), -0.5 * dt2);
  p.JVa = add(p.JVa, dR, -dt);
  p.JVg = add(p.JVg, mul(dRsa, p.JRg), -dt);
  const V3 wdt = scalev(w, dt);
  const M3 dRk = Exp(wdt), Jr = JrSO3(wdt);
  p.dR = nearest_rotation(mul(dR, dRk));
  const M3 dRkT = tr(dRk);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[9 * r + c] = dRkT(r, c);
      C[3 * r + c] = dt * Jr(r, c);
    }
  // cov = A cov A^T + B acc_noise B^T + C gyro_noise C^T (the noise covariances are isotropic)
  double T[81] = {}, N[81] = {};
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      double s = 0;
      for (int m = 0; m < 9; ++m) s += A[9 * i + m] * p.cov[9 * m + j];
      T[9 * i + j] = s;
    }
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      double s = 0;
      for (int m = 0; m < 9; ++m) s += T[9 * i + m] * A[9 * j + m];
      double bb = 0, cc = 0;
      for (int m = 0; m < 3; ++m) {
        bb += B[3 * i + m] * B[3 * j + m];
        cc += C[3 * i + m] * C[3 * j + m];
      }
      N[9 * i + j] = s + k.acc_noise * bb + k.gyro_noise * cc;
    }
  std::memcpy(p.cov, N, sizeof(N));
  for (int i = 0; i < 3; ++i) {
    p.cov_arw[4 * i] += k.acc_rw;
    p.cov_grw[4 * i] += k.gyro_rw;
  }
  p.JRg = add(mul(dRkT, p.JRg), Jr, -dt);
  p.dT += dt;
}

// Initialize + IntegrateNewMeasurement of every sample (imu_preintegration.cpp:117-150): the first step is
// 1/frequency, later ones the sample time difference.
void preint_integrate(Preint& p, const Calib& k, const sba_imu_measurement* m, int n, const V3& bg, const V3& ba) {
  preint_init(p, bg, ba);
  for (int i = 0; i < n; ++i) {
    const double dt = i == 0 ? 1.0 / k.freq : double(m[i].time_ns - m[i - 1].time_ns) * 1e-9;
    const V3 acc{{m[i].linear_acceleration[0], m[i].linear_acceleration[1], m[i].linear_acceleration[2]}};
    const V3 gyr{{m[i].angular_velocity[0], m[i].angular_velocity[1], m[i].angular_velocity[2]}};
    preint_update(p, k, acc, gyr, dt);
  }
}

// Reintegrate (imu_preintegration.cpp:129-143): integrate again at the updated biases.
void preint_reintegrate(Preint& p, const Calib& k, const sba_imu_measurement* all) {
  const V3 bg = p.bgu, ba = p.bau;
  const int first = p.first, n = p.n;
  preint_integrate(p, k, all + first, n, bg, ba);
  p.first = first;
  p.n = n;
}

// Inverse of a symmetric positive definite n x n matrix by Cholesky; false (and zero) when it is not.
bool spd_inverse(int n, const double* A, double* X) {
  std::vector<double> L(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = A[n * j + j];
    for (int k = 0; k < j; ++k) d -= L[n * j + k] * L[n * j + k];
    if (!(d > 0.0)) {
      std::fill(X, X + n * n, 0.0);
      return false;
    }
    const double l = std::sqrt(d);
    L[n * j + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double s = A[n * i + j];
      for (int k = 0; k < j; ++k) s -= L[n * i + k] * L[n * j + k];
      L[n * i + j] = s / l;
    }
  }
  for (int c = 0; c < n; ++c) {  // solve L L^T x = e_c
    std::vector<double> y(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[n * i + k] * y[k];
      y[i] = s / L[n * i + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= L[n * k + i] * X[n * k + c];
      X[n * i + c] = s / L[n * i + i];
    }
  }
  for (int i = 0; i < n; ++i)  // symmetric to the last bit
    for (int j = 0; j < i; ++j) X[n * i + j] = X[n * j + i] = 0.5 * (X[n * i + j] + X[n * j + i]);
  return true;
}

// InfoMatrix, InfoGyroRWMatrix, InfoAccRWMatrix (imu_preintegration.cpp:190-240): zero when empty; otherwise the
// inverse (the reference's FP32 QR solve with its invertibility test, then eigenvalues below 1e-12 clipped; the
// inverse of a positive definite covariance has none, so the clip changes nothing).
void preint_info(const Preint& p, double* info, double* igrw, double* iarw) {
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

// GetDeltaRotation / GetDeltaVelocity / GetDeltaPosition at biases (bg, ba) (imu_preintegration.cpp:162-181).
M3 delta_R(const Preint& p, const V3& bg) { return nearest_rotation(mul(p.dR, Exp(mul(p.JRg, addv(bg, p.bg, -1.0))))); }
V3 delta_V(const Preint& p, const V3& bg, const V3& ba) {
  return addv(addv(p.dV, mul(p.JVg, addv(bg, p.bg, -1.0))), mul(p.JVa, addv(ba, p.ba, -1.0)));
}
V3 delta_P(const Preint& p, const V3& bg, const V3& ba) {
  return addv(addv(p.dP, mul(p.JPg, addv(bg, p.bg, -1.0))), mul(p.JPa, addv(ba, p.ba, -1.0)));
}

void preint_from_abi(Preint& p, const sba_imu_preintegration& s) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) p.cov[9 * i + j] = s.cov[9 * j + i];
  auto m3 = [](M3& m, const float* a) { for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) m(r, c) = a[3 * c + r]; };
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      p.cov_arw[3 * r + c] = s.acc_rw_cov[3 * c + r];
      p.cov_grw[3 * r + c] = s.gyro_rw_cov[3 * c + r];
    }
  m3(p.dR, s.dR); m3(p.JRg, s.JRg); m3(p.JVg, s.JVg); m3(p.JVa, s.JVa); m3(p.JPg, s.JPg); m3(p.JPa, s.JPa);
  for (int i = 0; i < 3; ++i) {
    p.dV[i] = s.dV[i]; p.dP[i] = s.dP[i];
    p.bg[i] = s.gyro_bias[i]; p.ba[i] = s.acc_bias[i];
    p.bgu[i] = s.gyro_bias_updated[i]; p.bau[i] = s.acc_bias_updated[i];
  }
  p.dT = s.dT_s;
  p.first = s.first_measurement;
  p.n = s.num_measurements;
}

void preint_to_abi(const Preint& p, sba_imu_preintegration& s) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) s.cov[9 * j + i] = float(p.cov[9 * i + j]);
  auto m3 = [](const M3& m, float* a) { for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) a[3 * c + r] = float(m(r, c)); };
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      s.acc_rw_cov[3 * c + r] = float(p.cov_arw[3 * r + c]);
      s.gyro_rw_cov[3 * c + r] = float(p.cov_grw[3 * r + c]);
    }
  m3(p.dR, s.dR); m3(p.JRg, s.JRg); m3(p.JVg, s.JVg); m3(p.JVa, s.JVa); m3(p.JPg, s.JPg); m3(p.JPa, s.JPa);
  for (int i = 0; i < 3; ++i) {
    s.dV[i] = float(p.dV[i]); s.dP[i] = float(p.dP[i]);
    s.gyro_bias[i] = float(p.bg[i]); s.acc_bias[i] = float(p.ba[i]);
    s.gyro_bias_updated[i] = float(p.bgu[i]); s.acc_bias_updated[i] = float(p.bau[i]);
  }
  s.dT_s = float(p.dT);
  s.first_measurement = p.first;
  s.num_measurements = p.n;
}

// ---- the solver ----------------------------------------------------------------------------------------------------

struct Pose {
  M3 R;  // w_from_imu
  V3 t, v, bg, ba;
};

struct Edge {  // the inertial term between key frames i and i+1 (model of imu_sba_fixed_vel.cpp:575-651)
  double r[9], rg[3], ra[3];
  double J[8][27];  // 9x3 row-major: JR_left, Jt_left, Jv_left, Jbg_left, Jba_left, JR_right, Jt_right, Jv_right
  double info[81], igrw[9], iarw[9];  // unscaled
  bool active;
};

struct Obs {  // model of one observation (imu_sba_fixed_vel.cpp:499-566)
  double r[2], w;
  double JR[6], Jt[6], Jp[6];  // 2x3 row-major
  bool skipped;
};

struct Problem {
  int P, N, K, F, C;
  std::vector<double> pts;
  std::vector<double> xy, info;  // info row-major 2x2
  std::vector<int> pid, kid, cid;
  std::vector<M3> Rci;  // cam_from_imu
  std::vector<V3> tci;
  std::vector<Pose> poses;
  std::vector<Preint> pre;
  const sba_imu_measurement* meas;
  V3 g;
  int max_iterations;
  double scale, prior_gyro, prior_acc, imu_penalty, acc_rw_penalty, boundary_penalty, reint;
  Calib calib;
};

bool is_fixed(const Problem& p, int i) { return i < p.F; }
double huber(double s, double d) {  // math/robust_cost_function.cpp:62-74
  if (d <= 0) return 0.5 * s;
  return s < d * d ? 0.5 * s : d * std::sqrt(s) - 0.5 * d * d;
}
double dhuber(double s, double d) {  // robust_cost_function.cpp:76-87
  if (d <= 0) return 0.5;
  return s < d * d ? 0.5 : 0.5 * d / std::sqrt(s);
}
double edge_penalty(const Problem& p, int i) {  // boundary edge: i fixed, i+1 free
  return (is_fixed(p, i) && !is_fixed(p, i + 1)) ? p.boundary_penalty : p.imu_penalty;
}
double acc_rw_pen(const Problem& p) { return p.acc_rw_penalty >= 0 ? p.acc_rw_penalty : p.imu_penalty; }

void model_obs(const Problem& p, const std::vector<double>& pts, const std::vector<Pose>& poses, std::vector<Obs>& out,
               int* skipped) {
  out.assign(p.N, Obs{});
  int sk = 0;
  for (int o = 0; o < p.N; ++o) {
    Obs& m = out[o];
    const Pose& X = poses[p.kid[o]];
    const V3 pw{{pts[3 * p.pid[o]], pts[3 * p.pid[o] + 1], pts[3 * p.pid[o] + 2]}};
    const M3 RT = tr(X.R);
    const V3 pi = mul(RT, addv(pw, X.t, -1.0));
    const M3& Rc = p.Rci[p.cid[o]];
    const V3 pc = addv(mul(Rc, pi), p.tci[p.cid[o]]);
    if (!(pc[2] > 0)) {  // imu_sba_fixed_vel.cpp:518
      m.skipped = true;
      ++sk;
      continue;
    }
    const double iz = 1.0 / pc[2], u = pc[0] * iz, v = pc[1] * iz;
    m.r[0] = u - p.xy[2 * o];
    m.r[1] = v - p.xy[2 * o + 1];
    double M[6];  // dproj * R_ci
    for (int c = 0; c < 3; ++c) {
      M[c] = iz * Rc(0, c) - u * iz * Rc(2, c);
      M[3 + c] = iz * Rc(1, c) - v * iz * Rc(2, c);
    }
    const M3 S = skew(pi);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) {
        double jr = 0, jp = 0;
        for (int k = 0; k < 3; ++k) {
          jr += M[3 * r + k] * S(k, c);
          jp += M[3 * r + k] * RT(k, c);
        }
        m.JR[3 * r + c] = jr;
        m.Jp[3 * r + c] = jp;
        m.Jt[3 * r + c] = -M[3 * r + c];
      }
    const double* L = &p.info[4 * o];
    const double s = m.r[0] * (L[0] * m.r[0] + L[1] * m.r[1]) + m.r[1] * (L[2] * m.r[0] + L[3] * m.r[1]);
    m.w = dhuber(s, p.scale);
  }
  if (skipped) *skipped = sk;
}

// Inertial residual and Jacobians of edge i (imu_sba_fixed_vel.cpp:575-651).
void model_edge(const Problem& p, const std::vector<Pose>& poses, int i, Edge& e) {
  e.active = !(is_fixed(p, i) && is_fixed(p, i + 1));
  if (!e.active) return;
  const Preint& pr = p.pre[i];
  const Pose& A = poses[i];
  const Pose& B = poses[i + 1];
  const M3 R1T = tr(A.R), R2 = B.R;
  const M3 dR = delta_R(pr, A.bg);
  const V3 dV = delta_V(pr, A.bg, A.ba), dP = delta_P(pr, A.bg, A.ba);
  const double dT = pr.dT;
  const V3 er = Log(mul(mul(tr(dR), R1T), R2));
  const V3 vt = mul(R1T, addv(addv(B.v, A.v, -1.0), p.g, -dT));
  const V3 pt = mul(R1T, addv(addv(addv(B.t, A.t, -1.0), A.v, -dT), p.g, -0.5 * dT * dT));
  for (int k = 0; k < 3; ++k) {
    e.r[k] = er[k];
    e.r[3 + k] = vt[k] - dV[k];
    e.r[6 + k] = pt[k] - dP[k];
    e.rg[k] = A.bg[k] - B.bg[k];
    e.ra[k] = A.ba[k] - B.ba[k];
  }
  for (auto& J : e.J) std::fill(J, J + 27, 0.0);
  auto put = [&](int j, int row0, const M3& m) {
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) e.J[j][3 * (row0 + r) + c] = m(r, c);
  };
  const M3 JrInv = JrInvSO3(er);
  put(0, 0, scale(mul(mul(JrInv, tr(R2)), A.R), -1.0));
  put(0, 3, skew(vt));
  put(0, 6, skew(pt));
  put(1, 6, scale(eye3(), -1.0));
  put(2, 3, scale(R1T, -1.0));
  put(2, 6, scale(R1T, -dT));
  put(4, 3, scale(pr.JVa, -1.0));
  put(4, 6, scale(pr.JPa, -1.0));
  const V3 dbg = addv(pr.bgu, pr.bg, -1.0);  // GetDeltaBias: the stored updated-minus-original difference
  put(3, 0, scale(mul(mul(JlInvSO3(er), JrSO3(mul(pr.JRg, dbg))), pr.JRg), -1.0));
  put(3, 3, scale(pr.JVg, -1.0));
  put(3, 6, scale(pr.JPg, -1.0));
  put(5, 0, JrInv);
  put(6, 6, mul(R1T, R2));
  put(7, 3, R1T);
  preint_info(pr, e.info, e.igrw, e.iarw);
}

// Cost of a state (sba_imu_v1.cu:416-748): Huber reprojection terms (skipped at z <= 0), inertial and random-walk
// terms, bias priors on the first free key frame. Infinity when every observation is skipped.
double cost_of(const Problem& p, const std::vector<double>& pts, const std::vector<Pose>& poses) {
  double c = 0;
  int sk = 0;
  for (int o = 0; o < p.N; ++o) {
    const Pose& X = poses[p.kid[o]];
    const V3 pw{{pts[3 * p.pid[o]], pts[3 * p.pid[o] + 1], pts[3 * p.pid[o] + 2]}};
    const V3 pi = mul(tr(X.R), addv(pw, X.t, -1.0));
    const V3 pc = addv(mul(p.Rci[p.cid[o]], pi), p.tci[p.cid[o]]);
    if (!(pc[2] > 0)) {
      ++sk;
      continue;
    }
    const double r0 = pc[0] / pc[2] - p.xy[2 * o], r1 = pc[1] / pc[2] - p.xy[2 * o + 1];
    const double* L = &p.info[4 * o];
    c += huber(r0 * (L[0] * r0 + L[1] * r1) + r1 * (L[2] * r0 + L[3] * r1), p.scale);
  }
  if (sk == p.N) return std::numeric_limits<double>::infinity();
  for (int i = 0; i + 1 < p.K; ++i) {
    Edge e;
    model_edge(p, poses, i, e);
    if (!e.active) continue;
    const double pen = edge_penalty(p, i);
    double q = 0, qg = 0, qa = 0;
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) q += e.r[a] * e.info[9 * a + b] * e.r[b];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        qg += e.rg[a] * e.igrw[3 * a + b] * e.rg[b];
        qa += e.ra[a] * e.iarw[3 * a + b] * e.ra[b];
      }
    c += pen * qg + acc_rw_pen(p) * qa + pen * q;
  }
  if (p.F < p.K) {
    const Pose& X = poses[p.F];
    c += p.prior_gyro * dot(X.bg, X.bg) + p.prior_acc * dot(X.ba, X.ba);
  }
  return c;
}

struct System {  // full system (imu_sba_fixed_vel.cpp:654-1004)
  int n;                     // free key frames
  std::vector<double> U;     // 15n x 15n, row-major
  std::vector<double> ea;    // 15n
  std::vector<double> V;     // P x 9
  std::vector<double> eb;    // 3P
  std::vector<double> W;     // 3P x 6n (point rows, visual pose columns), dense
};

void build_full(const Problem& p, const std::vector<Pose>& poses, const std::vector<Obs>& obs,
                const std::vector<Edge>& edges, System& S) {
  const int n = p.K - p.F, N15 = 15 * n, N6 = 6 * n;
  S.n = n;
  S.U.assign(size_t(N15) * N15, 0.0);
  S.ea.assign(N15, 0.0);
  S.V.assign(size_t(p.P) * 9, 0.0);
  S.eb.assign(size_t(3) * p.P, 0.0);
  S.W.assign(size_t(3) * p.P * N6, 0.0);
  auto U = [&](int r, int c) -> double& { return S.U[size_t(r) * N15 + c]; };
  for (int o = 0; o < p.N; ++o) {
    const Obs& m = obs[o];
    if (m.skipped) continue;
    const double* L = &p.info[4 * o];
    const int i = p.pid[o], k = p.kid[o], id = k - p.F;
    // J = [JR Jt | Jp] (2x9); blocks w J^T L J and -J^T L w r
    double J[2][9];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) {
        J[r][c] = m.JR[3 * r + c];
        J[r][3 + c] = m.Jt[3 * r + c];
        J[r][6 + c] = m.Jp[3 * r + c];
      }
    double LJ[2][9], Lr[2];
    for (int r = 0; r < 2; ++r) {
      for (int c = 0; c < 9; ++c) LJ[r][c] = L[2 * r] * J[0][c] + L[2 * r + 1] * J[1][c];
      Lr[r] = L[2 * r] * m.r[0] + L[2 * r + 1] * m.r[1];
    }
    double H[9][9], g[9];
    for (int a = 0; a < 9; ++a) {
      for (int b = 0; b < 9; ++b) H[a][b] = m.w * (J[0][a] * LJ[0][b] + J[1][a] * LJ[1][b]);
      g[a] = -m.w * (J[0][a] * Lr[0] + J[1][a] * Lr[1]);
    }
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) S.V[9 * i + 3 * a + b] += H[6 + a][6 + b];
      S.eb[3 * i + a] += g[6 + a];
    }
    if (id >= 0) {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 6; ++b) S.W[(size_t(3) * i + a) * N6 + 6 * id + b] += H[6 + a][b];
      for (int a = 0; a < 6; ++a) {
        for (int b = 0; b < 6; ++b) U(15 * id + a, 15 * id + b) += H[a][b];
        S.ea[15 * id + a] += g[a];
      }
    }
  }
  // inertial and random-walk terms: edge i couples the 15-dof states of key frames i and i+1; the variables of its
  // Jacobian J (9 x 24) are [r1 t1 v1 bg1 ba1 | r2 t2 v2]
  const double apen = acc_rw_pen(p);
  for (int i = 0; i + 1 < p.K; ++i) {
    const Edge& e = edges[i];
    if (!e.active) continue;
    const double pen = edge_penalty(p, i);
    double J[9][24];
    for (int j = 0; j < 8; ++j)
      for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 3; ++c) J[r][3 * j + c] = e.J[j][3 * r + c];
    double LJ[9][24], Lr[9];
    for (int r = 0; r < 9; ++r) {
      Lr[r] = 0;
      for (int m = 0; m < 9; ++m) Lr[r] += pen * e.info[9 * r + m] * e.r[m];
      for (int c = 0; c < 24; ++c) {
        double s = 0;
        for (int m = 0; m < 9; ++m) s += pen * e.info[9 * r + m] * J[m][c];
        LJ[r][c] = s;
      }
    }
    const int id1 = i - p.F, id2 = i + 1 - p.F;
    auto col = [&](int v) { return v < 15 ? 15 * id1 + v : 15 * id2 + (v - 15); };  // variable -> system index
    auto free_var = [&](int v) { return v < 15 ? !is_fixed(p, i) : !is_fixed(p, i + 1); };
    for (int a = 0; a < 24; ++a) {
      if (!free_var(a)) continue;
      for (int b = 0; b < 24; ++b) {
        if (!free_var(b)) continue;
        double s = 0;
        for (int r = 0; r < 9; ++r) s += J[r][a] * LJ[r][b];
        U(col(a), col(b)) += s;
      }
      double s = 0;
      for (int r = 0; r < 9; ++r) s += J[r][a] * Lr[r];
      S.ea[col(a)] -= s;
    }
    // random walks r = b_i - b_{i+1}: d r / d b_i = I, d r / d b_{i+1} = -I
    for (int which = 0; which < 2; ++which) {
      const double* I3 = which == 0 ? e.igrw : e.iarw;
      const double* r = which == 0 ? e.rg : e.ra;
      const double pw = which == 0 ? pen : apen;
      const int off = which == 0 ? 9 : 12;
      double Ir[3];
      for (int a = 0; a < 3; ++a) Ir[a] = pw * (I3[3 * a] * r[0] + I3[3 * a + 1] * r[1] + I3[3 * a + 2] * r[2]);
      const bool f1 = !is_fixed(p, i), f2 = !is_fixed(p, i + 1);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double v = pw * I3[3 * a + b];
          if (f1) U(15 * id1 + off + a, 15 * id1 + off + b) += v;
          if (f2) U(15 * id2 + off + a, 15 * id2 + off + b) += v;
          if (f1 && f2) {
            U(15 * id1 + off + a, 15 * id2 + off + b) -= v;
            U(15 * id2 + off + a, 15 * id1 + off + b) -= v;
          }
        }
      for (int a = 0; a < 3; ++a) {
        if (f1) S.ea[15 * id1 + off + a] -= Ir[a];
        if (f2) S.ea[15 * id2 + off + a] += Ir[a];
      }
    }
  }
  if (n > 0) {  // bias priors on the first free key frame (imu_sba_fixed_vel.cpp:992-1003)
    const Pose& X = poses[p.F];
    for (int a = 0; a < 3; ++a) {
      U(9 + a, 9 + a) += p.prior_gyro;
      U(12 + a, 12 + a) += p.prior_acc;
      S.ea[9 + a] -= p.prior_gyro * X.bg[a];
      S.ea[12 + a] -= p.prior_acc * X.ba[a];
    }
  }
}

// Inverse of a damped 3x3 point block by the reference's pivoted LDLT (cuda_sba_common.h:365-470; used by
// sba_imu_v1.cu:268-276): the largest diagonal first, then the larger remaining diagonal (compared before its
// update); negative pivots are clipped to zero and a zero pivot contributes nothing (its D^-1 entry is 0), so the
// block of a point without observations inverts to zero.
void ldlt_inverse3(const double* A, double* X) {
  double a00 = A[0], a11 = A[4], a22 = A[8], a10 = A[3], a20 = A[6], a21 = A[7];
  bool p01 = false, p02 = false, p12 = false;
  const double d0 = std::fabs(a00), d1 = std::fabs(a11), d2 = std::fabs(a22);
  if (std::max(d1, d2) > d0) {
    if (d1 > d2) {
      std::swap(a20, a21);
      std::swap(a00, a11);
      p01 = true;
    } else {
      std::swap(a00, a22);
      std::swap(a10, a21);
      p02 = true;
    }
  }
  a00 = std::max(a00, 0.0);
  a10 = a00 == 0.0 ? 0.0 : a10 / a00;
  a20 = a00 == 0.0 ? 0.0 : a20 / a00;
  if (std::fabs(a22) > std::fabs(a11)) {
    std::swap(a10, a20);
    std::swap(a11, a22);
    p12 = true;
  }
  a11 = std::max(a11 - a10 * a10 * a00, 0.0);
  a21 = a11 == 0.0 ? 0.0 : (a21 - a20 * a10 * a00) / a11;
  a22 = std::max(a22 - a20 * a20 * a00 - a21 * a21 * a11, 0.0);
  const double i0 = a00 == 0.0 ? 0.0 : 1.0 / a00, i1 = a11 == 0.0 ? 0.0 : 1.0 / a11, i2 = a22 == 0.0 ? 0.0 : 1.0 / a22;
  for (int c = 0; c < 3; ++c) {
    double b[3] = {double(c == 0), double(c == 1), double(c == 2)};
    if (p01) std::swap(b[0], b[1]);
    if (p02) std::swap(b[0], b[2]);
    if (p12) std::swap(b[1], b[2]);
    const double y0 = b[0], y1 = b[1] - y0 * a10, y2 = b[2] - y0 * a20 - y1 * a21;
    const double z0 = y0 * i0, z1 = y1 * i1, z2 = y2 * i2;
    double x[3];
    x[2] = z2;
    x[1] = z1 - x[2] * a21;
    x[0] = z0 - x[1] * a10 - x[2] * a20;
    if (p12) std::swap(x[1], x[2]);
    if (p02) std::swap(x[0], x[2]);
    if (p01) std::swap(x[0], x[1]);
    X[c] = x[0]; X[3 + c] = x[1]; X[6 + c] = x[2];
  }
}

// Dense Cholesky solve of the reduced system (lower); false on a non-positive pivot.
bool chol_solve(int n, std::vector<double>& A, std::vector<double>& b) {
  for (int j = 0; j < n; ++j) {
    double d = A[size_t(n) * j + j];
    for (int k = 0; k < j; ++k) d -= A[size_t(n) * j + k] * A[size_t(n) * j + k];
    if (!(d > 0.0)) return false;
    const double l = std::sqrt(d);
    A[size_t(n) * j + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double s = A[size_t(n) * i + j];
      for (int k = 0; k < j; ++k) s -= A[size_t(n) * i + k] * A[size_t(n) * j + k];
      A[size_t(n) * i + j] = s / l;
    }
  }
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[size_t(n) * i + k] * b[k];
    b[i] = s / A[size_t(n) * i + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= A[size_t(n) * k + i] * b[k];
    b[i] = s / A[size_t(n) * i + i];
  }
  return true;
}

struct Step {
  std::vector<double> dc, dp;  // 15n, 3P
  bool ok;
};

// Reduced system, its solve and the back-substitution (imu_sba_fixed_vel.cpp:1006-1071, 338-374; the pose system is
// factored by Cholesky as the GPU bundler does).
void compute_step(const Problem& p, const System& S, float lambda, Step& st, std::vector<double>* reduced = nullptr,
                  std::vector<double>* reduced_rhs = nullptr) {
  const int n = S.n, N15 = 15 * n, N6 = 6 * n;
  const double lam = lambda;
  std::vector<double> Vinv(size_t(9) * p.P), r(size_t(3) * p.P);
  for (int i = 0; i < p.P; ++i) {
    double m[9];
    for (int k = 0; k < 9; ++k) m[k] = S.V[9 * i + k];
    for (int k = 0; k < 3; ++k) m[4 * k] += lam * S.V[9 * i + 4 * k];
    ldlt_inverse3(m, &Vinv[9 * i]);
    for (int a = 0; a < 3; ++a)
      r[3 * i + a] = Vinv[9 * i + 3 * a] * S.eb[3 * i] + Vinv[9 * i + 3 * a + 1] * S.eb[3 * i + 1] +
                     Vinv[9 * i + 3 * a + 2] * S.eb[3 * i + 2];
  }
  std::vector<double> A = S.U, b = S.ea;
  // S_jk -= sum_i W_ij^T Vinv_i W_ik; rhs_j -= sum_i W_ij^T Vinv_i e_b,i  (only the 6 visual dims of each pose)
  std::vector<int> nz;
  for (int i = 0; i < p.P; ++i) {
    nz.clear();
    for (int c = 0; c < N6; c += 6) {
      bool any = false;
      for (int a = 0; a < 3 && !any; ++a)
        for (int q = 0; q < 6; ++q) any |= S.W[(size_t(3) * i + a) * N6 + c + q] != 0.0;
      if (any) nz.push_back(c / 6);
    }
    for (int j : nz) {
      double Y[6][3];  // W_ij^T Vinv_i
      for (int q = 0; q < 6; ++q)
        for (int a = 0; a < 3; ++a) {
          double s = 0;
          for (int m = 0; m < 3; ++m) s += S.W[(size_t(3) * i + m) * N6 + 6 * j + q] * Vinv[9 * i + 3 * m + a];
          Y[q][a] = s;
        }
      for (int q = 0; q < 6; ++q)
        b[15 * j + q] -= Y[q][0] * S.eb[3 * i] + Y[q][1] * S.eb[3 * i + 1] + Y[q][2] * S.eb[3 * i + 2];
      for (int k : nz)
        for (int q = 0; q < 6; ++q)
          for (int s2 = 0; s2 < 6; ++s2) {
            double s = 0;
            for (int a = 0; a < 3; ++a) s += Y[q][a] * S.W[(size_t(3) * i + a) * N6 + 6 * k + s2];
            A[size_t(15 * j + q) * N15 + 15 * k + s2] -= s;
          }
    }
  }
  for (int d = 0; d < N15; ++d) A[size_t(d) * N15 + d] += lam * S.U[size_t(d) * N15 + d];
  if (reduced) *reduced = A;
  if (reduced_rhs) *reduced_rhs = b;
  st.ok = chol_solve(N15, A, b);
  st.dc = b;
  st.dp.assign(size_t(3) * p.P, 0.0);
  if (!st.ok) return;
  for (int i = 0; i < p.P; ++i) {  // dp_i = Vinv_i (e_b,i - sum_j W_ij dc_j)
    double t[3] = {S.eb[3 * i], S.eb[3 * i + 1], S.eb[3 * i + 2]};
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < N6; ++c) {
        const double w = S.W[(size_t(3) * i + a) * N6 + c];
        if (w != 0.0) t[a] -= w * b[15 * (c / 6) + c % 6];
      }
    for (int a = 0; a < 3; ++a)
      st.dp[3 * i + a] = Vinv[9 * i + 3 * a] * t[0] + Vinv[9 * i + 3 * a + 1] * t[1] + Vinv[9 * i + 3 * a + 2] * t[2];
  }
}

// The state after a step (imu_sba_fixed_vel.cpp:224-240): R <- R Exp(dr), t <- t + R dt, v, b += ...
void apply_step(const Problem& p, const Step& st, const std::vector<double>& pts, const std::vector<Pose>& poses,
                std::vector<double>& pts2, std::vector<Pose>& poses2) {
  pts2 = pts;
  for (size_t k = 0; k < pts2.size(); ++k) pts2[k] += st.dp[k];
  poses2 = poses;
  for (int k = p.F; k < p.K; ++k) {
    const double* d = &st.dc[15 * (k - p.F)];
    Pose& X = poses2[k];
    const V3 dr{{d[0], d[1], d[2]}}, dt{{d[3], d[4], d[5]}};
    X.t = addv(X.t, mul(X.R, dt));
    X.R = mul(X.R, Exp(dr));
    for (int a = 0; a < 3; ++a) {
      X.v[a] += d[6 + a];
      X.bg[a] += d[9 + a];
      X.ba[a] += d[12 + a];
    }
  }
}

// d^T (H + 2 lambda diag H) d over the full system, and whether any |dp|, |dc| reaches FLT_EPSILON
// (sba_imu_v1.cu:767-852).
double predicted(const Problem& p, const System& S, const Step& st, float lambda, bool* sig_p, bool* sig_c) {
  const int n = S.n, N15 = 15 * n, N6 = 6 * n;
  const double l2 = 1.0 + 2.0 * double(lambda);
  double pred = 0, mp = 0, mc = 0;
  for (int i = 0; i < p.P; ++i) {
    const double* dp = &st.dp[3 * i];
    for (int a = 0; a < 3; ++a) {
      mp = std::max(mp, std::fabs(dp[a]));
      for (int b = 0; b < 3; ++b) pred += dp[a] * S.V[9 * i + 3 * a + b] * dp[b] * (a == b ? l2 : 1.0);
      double v = 0;
      for (int c = 0; c < N6; ++c) v += S.W[(size_t(3) * i + a) * N6 + c] * st.dc[15 * (c / 6) + c % 6];
      pred += 2.0 * dp[a] * v;
    }
  }
  for (int r = 0; r < N15; ++r) {
    mc = std::max(mc, std::fabs(st.dc[r]));
    for (int c = 0; c < N15; ++c) pred += st.dc[r] * st.dc[c] * S.U[size_t(r) * N15 + c] * (r == c ? l2 : 1.0);
  }
  *sig_p = mp >= kFltEps;
  *sig_c = mc >= kFltEps;
  return pred;
}

void build_model(const Problem& p, const std::vector<double>& pts, const std::vector<Pose>& poses,
                 std::vector<Obs>& obs, std::vector<Edge>& edges, System& S) {
  model_obs(p, pts, poses, obs, nullptr);
  edges.assign(std::max(p.K - 1, 0), Edge{});
  for (int i = 0; i + 1 < p.K; ++i) model_edge(p, poses, i, edges[i]);
  build_full(p, poses, obs, edges, S);
}

bool load(Problem& p, const sba_imu_calibration* calib, const sba_imu_problem_view* v) {
  p.P = v->num_points;
  p.N = v->num_observations;
  p.K = v->num_poses;
  p.F = std::min(std::max(v->num_fixed_key_frames, 0), v->num_poses);
  p.C = v->num_cameras;
  p.calib = calib_of(*calib);
  p.pts.assign(v->points, v->points + 3 * size_t(p.P));
  p.xy.assign(v->observation_xys, v->observation_xys + 2 * size_t(p.N));
  p.info.resize(4 * size_t(p.N));
  for (int o = 0; o < p.N; ++o) {  // column-major in, row-major here
    const float* L = v->observation_infos + 4 * o;
    p.info[4 * o] = L[0]; p.info[4 * o + 1] = L[2]; p.info[4 * o + 2] = L[1]; p.info[4 * o + 3] = L[3];
  }
  p.pid.assign(v->point_ids, v->point_ids + p.N);
  p.kid.assign(v->pose_ids, v->pose_ids + p.N);
  p.cid.assign(v->camera_ids, v->camera_ids + p.N);
  for (int c = 0; c < p.C; ++c) {  // cam_from_imu = camera_from_rig * rig_from_imu
    const float* T = v->camera_from_rig + 16 * c;
    M3 R{};
    V3 t{};
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) R(r, k) = T[4 * k + r];
      t[r] = T[12 + r];
    }
    p.Rci.push_back(mul(R, p.calib.R_ri));
    p.tci.push_back(addv(mul(R, p.calib.t_ri), t));
  }
  p.poses.resize(p.K);
  p.pre.resize(p.K);
  for (int k = 0; k < p.K; ++k) {
    const float* T = v->w_from_imu + 16 * k;
    Pose& X = p.poses[k];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) X.R(r, c) = T[4 * c + r];
      X.t[r] = T[12 + r];
      X.v[r] = v->velocity[3 * k + r];
      X.bg[r] = v->gyro_bias[3 * k + r];
      X.ba[r] = v->acc_bias[3 * k + r];
    }
    preint_from_abi(p.pre[k], v->preintegration[k]);
  }
  p.meas = v->measurements;
  for (int r = 0; r < 3; ++r) p.g[r] = v->gravity[r];
  p.max_iterations = v->max_iterations;
  p.scale = v->robustifier_scale;
  p.prior_gyro = v->prior_gyro;
  p.prior_acc = v->prior_acc;
  p.imu_penalty = v->imu_penalty;
  p.acc_rw_penalty = v->acc_rw_penalty;
  p.boundary_penalty = v->boundary_imu_penalty;
  p.reint = v->reintegration_thresh;
  return true;
}

// SetNewBias and, past the threshold, Reintegrate of every free key frame's preintegration
// (imu_sba_gpu_fixed_vel.cpp:286-301).
void reintegrate_moved(Problem& p, const std::vector<Pose>& poses, bool from_original) {
  for (int k = p.F; k < p.K; ++k) {
    Preint& pr = p.pre[k];
    const Pose& X = poses[k];
    const V3 dg = addv(X.bg, pr.bg, -1.0), da = addv(X.ba, pr.ba, -1.0);
    if (from_original) {  // the final check compares with the original biases first (:339-356)
      if (norm(dg) > p.reint || norm(da) > p.reint) {
        pr.bgu = X.bg;
        pr.bau = X.ba;
        preint_reintegrate(pr, p.calib, p.meas);
      }
      continue;
    }
    pr.bgu = X.bg;
    pr.bau = X.ba;
    if (norm(dg) > p.reint || norm(da) > p.reint) preint_reintegrate(pr, p.calib, p.meas);
  }
}

}  // namespace

extern "C" {

struct OrcImuResult {
  double* points;     // [P][3]
  double* poses;      // [K][12]: R row-major (9), t (3)
  double* vel_bias;   // [K][9]: v, bg, ba
  sba_imu_preintegration* preintegration;  // [K] (may be null)
  int32_t status, iterations, accepted, pad;
  double initial_cost, last_cost, lambda;
  int32_t trace_capacity, pad2;
  double* trace_lambda;
  double* trace_cost;
  double* trace_rho;
  int32_t* trace_flags;  // 1 accepted, 2 factorization failed, 4 stop test, 8 non-finite trial cost
};

// The GPU bundler's solve (imu_sba_gpu_fixed_vel.cpp:75-363) in FP64.
int orc_imu_solve(const sba_imu_calibration* calib, const sba_imu_problem_view* v, OrcImuResult* res) {
  Problem p;
  load(p, calib, v);
  res->iterations = 0;
  res->accepted = 0;
  res->lambda = 0;
  auto finish = [&](int status, const std::vector<double>& pts, const std::vector<Pose>& poses) {
    res->status = status;
    for (size_t k = 0; k < pts.size(); ++k) res->points[k] = pts[k];
    for (int k = 0; k < p.K; ++k) {
      for (int a = 0; a < 9; ++a) res->poses[12 * k + a] = poses[k].R.a[a];
      for (int a = 0; a < 3; ++a) {
        res->poses[12 * k + 9 + a] = poses[k].t[a];
        res->vel_bias[9 * k + a] = poses[k].v[a];
        res->vel_bias[9 * k + 3 + a] = poses[k].bg[a];
        res->vel_bias[9 * k + 6 + a] = poses[k].ba[a];
      }
      if (res->preintegration) preint_to_abi(p.pre[k], res->preintegration[k]);
    }
    return status;
  };
  std::vector<double> pts = p.pts;
  std::vector<Pose> poses = p.poses;
  if (p.K - 1 <= 0 || p.P == 0 || p.N == 0) return finish(SBA_E_EMPTY, pts, poses);
  const double initial = cost_of(p, pts, poses);
  res->initial_cost = initial;
  res->last_cost = initial;
  if (initial < kFltEps) return finish(SBA_OK, pts, poses);
  if (!std::isfinite(initial)) return finish(SBA_E_INFEASIBLE, pts, poses);
  std::vector<Obs> obs;
  std::vector<Edge> edges;
  System S;
  build_model(p, pts, poses, obs, edges, S);
  float lambda = 1e2f;
  double current = initial;
  int it = 0;
  auto trace = [&](int k, double rho, double c, int flags) {
    if (k < res->trace_capacity) {
      res->trace_lambda[k] = lambda;
      res->trace_cost[k] = c;
      res->trace_rho[k] = rho;
      res->trace_flags[k] = flags;
    }
  };
  while (it < p.max_iterations) {
    ++it;
    res->iterations = it;
    Step st;
    compute_step(p, S, lambda, st);
    if (!st.ok) {
      trace(it - 1, 0, 0, 2);
      lambda *= 5.f;
      continue;
    }
    std::vector<double> pts2;
    std::vector<Pose> poses2;
    apply_step(p, st, pts, poses, pts2, poses2);
    const double c = cost_of(p, pts2, poses2);
    bool sp = false, sc = false;
    const double pred = predicted(p, S, st, lambda, &sp, &sc) / current;
    if (!std::isfinite(c)) {
      trace(it - 1, 0, c, 8);
      lambda *= 5.f;
      continue;
    }
    if (current < initial * kFltEps || (pred < kFltEps && !sp && !sc)) {
      trace(it - 1, 0, c, 4);
      break;
    }
    const double rho = (1.0 - c / current) / pred;
    trace(it - 1, rho, c, rho > 0.25 ? 1 : 0);
    if (rho > 0.25) {
      if (rho > 0.75 && lambda * 0.125f > 0.f) lambda *= 0.125f;
      current = c;
      pts = pts2;
      poses = poses2;
      ++res->accepted;
      if (it == p.max_iterations) break;
      reintegrate_moved(p, poses, false);
      build_model(p, pts, poses, obs, edges, S);
    } else {
      lambda *= 5.f;
    }
  }
  if (current < initial) reintegrate_moved(p, poses, true);
  res->last_cost = current;
  res->lambda = lambda;
  return finish(current < initial ? SBA_OK : SBA_E_NO_PROGRESS, pts, poses);
}

int orc_imu_integrate(const sba_imu_calibration* calib, const sba_imu_measurement* m, int32_t n,
                      const float* bg, const float* ba, sba_imu_preintegration* out) {
  Preint p;
  const Calib k = calib_of(*calib);
  preint_integrate(p, k, m, n, V3{{bg[0], bg[1], bg[2]}}, V3{{ba[0], ba[1], ba[2]}});
  p.first = 0;
  p.n = n;
  preint_to_abi(p, *out);
  return 0;
}

// The same integration, every quantity in FP64 (row-major): dR 9, dV 3, dP 3, JRg JVg JVa JPg JPa 45, cov 81,
// cov_grw 9, cov_arw 9, dT 1 = 160 doubles.
int orc_imu_integrate_f64(const sba_imu_calibration* calib, const sba_imu_measurement* m, int32_t n,
                          const double* bg, const double* ba, double* out) {
  Preint p;
  const Calib k = calib_of(*calib);
  preint_integrate(p, k, m, n, V3{{bg[0], bg[1], bg[2]}}, V3{{ba[0], ba[1], ba[2]}});
  double* o = out;
  for (double x : p.dR.a) *o++ = x;
  for (double x : p.dV.v) *o++ = x;
  for (double x : p.dP.v) *o++ = x;
  for (const M3* m3 : {&p.JRg, &p.JVg, &p.JVa, &p.JPg, &p.JPa})
    for (double x : m3->a) *o++ = x;
  for (double x : p.cov) *o++ = x;
  for (double x : p.cov_grw) *o++ = x;
  for (double x : p.cov_arw) *o++ = x;
  *o++ = p.dT;
  return 0;
}

int orc_imu_information(const sba_imu_preintegration* s, double* info, double* igrw, double* iarw) {
  Preint p;
  preint_from_abi(p, *s);
  preint_info(p, info, igrw, iarw);
  return 0;
}

// The cost of a state given as an override of the problem's (points [P][3], poses [K][12] R row-major + t,
// vel_bias [K][9]); null pointers keep the problem's values.
double orc_imu_cost(const sba_imu_calibration* calib, const sba_imu_problem_view* v, const double* points,
                    const double* poses, const double* vel_bias) {
  Problem p;
  load(p, calib, v);
  std::vector<double> pts = p.pts;
  if (points) pts.assign(points, points + 3 * size_t(p.P));
  std::vector<Pose> X = p.poses;
  for (int k = 0; k < p.K; ++k) {
    if (poses)
      for (int a = 0; a < 9; ++a) X[k].R.a[a] = poses[12 * k + a];
    for (int a = 0; a < 3; ++a) {
      if (poses) X[k].t[a] = poses[12 * k + 9 + a];
      if (vel_bias) {
        X[k].v[a] = vel_bias[9 * k + a];
        X[k].bg[a] = vel_bias[9 * k + 3 + a];
        X[k].ba[a] = vel_bias[9 * k + 6 + a];
      }
    }
  }
  return cost_of(p, pts, X);
}

// The model of the problem's state: the inertial residuals [K-1][9 + 3 + 3] and Jacobians [K-1][8][9][3] (row-major),
// the reprojection residuals [N][2], weights [N] and Jacobians [N][3][2][3] (JR, Jt, Jp).
int orc_imu_model(const sba_imu_calibration* calib, const sba_imu_problem_view* v, double* edge_res, double* edge_jac,
                  double* obs_res, double* obs_w, double* obs_jac) {
  Problem p;
  load(p, calib, v);
  for (int i = 0; i + 1 < p.K; ++i) {
    Edge e{};
    model_edge(p, p.poses, i, e);
    for (int a = 0; a < 9; ++a) edge_res[15 * i + a] = e.active ? e.r[a] : 0.0;
    for (int a = 0; a < 3; ++a) {
      edge_res[15 * i + 9 + a] = e.active ? e.rg[a] : 0.0;
      edge_res[15 * i + 12 + a] = e.active ? e.ra[a] : 0.0;
    }
    for (int j = 0; j < 8; ++j)
      for (int a = 0; a < 27; ++a) edge_jac[216 * i + 27 * j + a] = e.active ? e.J[j][a] : 0.0;
  }
  std::vector<Obs> obs;
  model_obs(p, p.pts, p.poses, obs, nullptr);
  for (int o = 0; o < p.N; ++o) {
    obs_res[2 * o] = obs[o].r[0];
    obs_res[2 * o + 1] = obs[o].r[1];
    obs_w[o] = obs[o].w;
    for (int a = 0; a < 6; ++a) {
      obs_jac[18 * o + a] = obs[o].JR[a];
      obs_jac[18 * o + 6 + a] = obs[o].Jt[a];
      obs_jac[18 * o + 12 + a] = obs[o].Jp[a];
    }
  }
  return 0;
}

// The full system of the problem's state (U [15n][15n], e_a [15n], V [P][9], e_b [3P], W [3P][6n]) and, for a
// damping lambda, the reduced system (A [15n][15n] before factoring, rhs [15n]) and the step (dc [15n], dp [3P]).
// Returns 1 when the reduced system factored.
int orc_imu_system(const sba_imu_calibration* calib, const sba_imu_problem_view* v, float lambda, double* U,
                   double* ea, double* V, double* eb, double* W, double* A, double* rhs, double* dc, double* dp) {
  Problem p;
  load(p, calib, v);
  std::vector<Obs> obs;
  std::vector<Edge> edges;
  System S;
  build_model(p, p.pts, p.poses, obs, edges, S);
  std::copy(S.U.begin(), S.U.end(), U);
  std::copy(S.ea.begin(), S.ea.end(), ea);
  std::copy(S.V.begin(), S.V.end(), V);
  std::copy(S.eb.begin(), S.eb.end(), eb);
  std::copy(S.W.begin(), S.W.end(), W);
  Step st;
  std::vector<double> Ar, br;
  compute_step(p, S, lambda, st, &Ar, &br);
  std::copy(Ar.begin(), Ar.end(), A);
  std::copy(br.begin(), br.end(), rhs);
  std::copy(st.dc.begin(), st.dc.end(), dc);
  std::copy(st.dp.begin(), st.dp.end(), dp);
  return st.ok ? 1 : 0;
}

// Exp, Log, the left/right Jacobians and their inverses, and the nearest rotation, for the tests' pins.
void orc_so3(int which, const double* in, double* out) {
  M3 m{};
  V3 x{{in[0], in[1], in[2]}};
  switch (which) {
    case 0: m = Exp(x); break;
    case 1: { M3 R{}; for (int a = 0; a < 9; ++a) R.a[a] = in[a]; const V3 l = Log(R); for (int a = 0; a < 3; ++a) out[a] = l[a]; return; }
    case 2: m = JlSO3(x); break;
    case 3: m = JlInvSO3(x); break;
    case 4: m = JrSO3(x); break;
    case 5: m = JrInvSO3(x); break;
    case 6: { M3 R{}; for (int a = 0; a < 9; ++a) R.a[a] = in[a]; m = nearest_rotation(R); break; }
    default: return;
  }
  for (int a = 0; a < 9; ++a) out[a] = m.a[a];
}

}  // extern "C"
