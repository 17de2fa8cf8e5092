// HIP kernels of the inertial sliding-window SBA (one window, FP64 throughout).
//
// One LM iteration (imu_sba_gpu_fixed_vel.cpp:136-214) is:
//   k_imu_point_inv   damped point blocks V* = V + lambda diag V inverted by the reference's pivoted LDLT
//                     (cuda_sba_common.h:365-470), Vinv e_b and Y = Vinv W for each (point, free key frame) pair
//   k_imu_reduced     the lower 15x15 blocks of S = U + lambda diag U - sum_i W_ij^T Vinv_i W_ik (only the rotation
//                     and translation columns of W are non-zero), each block summed over its co-observing points in a
//                     fixed order, and the reduced right-hand side
//   k_imu_factor      dense Cholesky of S and the two triangular solves in one workgroup
//   k_imu_backsub     dp = Vinv (e_b - sum_j W_ij dc_j), trial points, the points' share of the prediction
//   k_imu_pose_trial  trial key frames: R Exp(dr), t + R dt, v + dv, biases + db (imu_sba_fixed_vel.cpp:224-240)
//   k_imu_pose_pred   dc^T (U + 2 lambda diag U) dc by rows
//   k_imu_cost        the trial cost: Huber reprojection terms, inertial and random-walk terms, bias prior
//   k_imu_reduce      the iteration's scalars, summed in a fixed order (deterministic)
// and after an accepted step k_imu_accept, then the model (k_imu_obs_model, k_imu_edge_model) and the full system
// (k_imu_point_full, k_imu_pose_full) of the new state (imu_sba_fixed_vel.cpp:455-1004).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "sba_imu_device.h"
#include "sba_tiles.h"

using namespace sba_imu_dev;

namespace {

// ---- 3x3 FP64 helpers (row-major) --------------------------------------------------------------------------------

__device__ __forceinline__ void mm3(const double* A, const double* B, double* C) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}
__device__ __forceinline__ void mtm3(const double* A, const double* B, double* C) {  // A^T B
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
}
__device__ __forceinline__ void mv3(const double* A, const double* x, double* y) {
  for (int r = 0; r < 3; ++r) y[r] = A[3 * r] * x[0] + A[3 * r + 1] * x[1] + A[3 * r + 2] * x[2];
}
__device__ __forceinline__ void mtv3(const double* A, const double* x, double* y) {  // A^T x
  for (int r = 0; r < 3; ++r) y[r] = A[r] * x[0] + A[3 + r] * x[1] + A[6 + r] * x[2];
}
__device__ __forceinline__ void skew3(const double* w, double* S) {
  S[0] = 0; S[1] = -w[2]; S[2] = w[1];
  S[3] = w[2]; S[4] = 0; S[5] = -w[0];
  S[6] = -w[1]; S[7] = w[0]; S[8] = 0;
}

// Rodrigues (math/twist.cpp:75-82).
__device__ void exp3(const double* w, double* R) {
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double S[9];
  skew3(w, S);
  if (th < 1e-12) {
    for (int i = 0; i < 9; ++i) R[i] = S[i] + (i % 4 == 0 ? 1.0 : 0.0);
    return;
  }
  double s, c;
  sincos(th, &s, &c);
  const double a = s / th, b = (1.0 - c) / (th * th);
  double S2[9];
  mm3(S, S, S2);
  for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0 ? 1.0 : 0.0) + a * S[i] + b * S2[i];
}

// SO(3) logarithm: angle atan2(|2 sin|, 2 cos); near pi the axis from the symmetric part (sba_imu_v1.cu:96-111).
__device__ void log3(const double* R, double* w) {
  const double r0 = R[7] - R[5], r1 = R[2] - R[6], r2 = R[3] - R[1];
  const double two_sin = sqrt(r0 * r0 + r1 * r1 + r2 * r2), two_cos = R[0] + R[4] + R[8] - 1.0;
  const double th = atan2(two_sin, two_cos);
  if (two_cos > -1.9) {
    const double s = th < 1e-6 ? 0.5 * (1.0 + th * th / 6.0) : th / two_sin;
    w[0] = s * r0; w[1] = s * r1; w[2] = s * r2;
    return;
  }
  double S[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[3 * r + c] = R[3 * r + c] + R[3 * c + r] - (r == c ? two_cos : 0.0);
  int k = 0;
  if (S[4] > S[0]) k = 1;
  if (S[8] > S[4 * k]) k = 2;
  double n0 = S[k], n1 = S[3 + k], n2 = S[6 + k];
  const double inv = 1.0 / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  n0 *= inv; n1 *= inv; n2 *= inv;
  if (n0 * r0 + n1 * r1 + n2 * r2 < 0) { n0 = -n0; n1 = -n1; n2 = -n2; }
  w[0] = th * n0; w[1] = th * n1; w[2] = th * n2;
}

// Left Jacobian of SO(3) and its inverse at x (math/twist.cpp:103-126); the right ones are those of -x.
__device__ void jl3(const double* x, double sign, double* J) {
  const double y[3] = {sign * x[0], sign * x[1], sign * x[2]};
  const double phi = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  double S[9];
  skew3(y, S);
  if (phi < 1e-8) {
    for (int i = 0; i < 9; ++i) J[i] = (i % 4 == 0 ? 1.0 : 0.0) + 0.5 * S[i];
    return;
  }
  const double s = sin(phi) / phi, k = (1.0 - cos(phi)) / phi, ip = 1.0 / phi;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      J[3 * r + c] = (r == c ? s : 0.0) + (1.0 - s) * y[r] * y[c] * ip * ip + k * S[3 * r + c] * ip;
}
__device__ void jlinv3(const double* x, double sign, double* J) {
  const double y[3] = {sign * x[0], sign * x[1], sign * x[2]};
  const double phi = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  double S[9];
  skew3(y, S);
  if (phi < 1e-8) {
    for (int i = 0; i < 9; ++i) J[i] = (i % 4 == 0 ? 1.0 : 0.0) - 0.5 * S[i];
    return;
  }
  const double h = 0.5 * phi, cot = h / tan(h), ip = 1.0 / phi;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      J[3 * r + c] = (r == c ? cot : 0.0) + (1.0 - cot) * y[r] * y[c] * ip * ip - h * S[3 * r + c] * ip;
}

// The orthogonal polar factor of a near-rotation (U V^T of its SVD, common/rotation_utils.h:36-50): Newton's
// iteration X <- (X + X^-T) / 2, which squares the distance to SO(3) at each step.
__device__ void polar3(double* X) {
  for (int it = 0; it < 5; ++it) {
    double C[9];
    C[0] = X[4] * X[8] - X[5] * X[7];
    C[1] = X[5] * X[6] - X[3] * X[8];
    C[2] = X[3] * X[7] - X[4] * X[6];
    C[3] = X[2] * X[7] - X[1] * X[8];
    C[4] = X[0] * X[8] - X[2] * X[6];
    C[5] = X[1] * X[6] - X[0] * X[7];
    C[6] = X[1] * X[5] - X[2] * X[4];
    C[7] = X[2] * X[3] - X[0] * X[5];
    C[8] = X[0] * X[4] - X[1] * X[3];
    const double det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
    const double id = 0.5 / det;
    for (int i = 0; i < 9; ++i) X[i] = 0.5 * X[i] + id * C[i];
  }
}

__device__ __forceinline__ double huber(double s, double d) {  // math/robust_cost_function.cpp:62-74
  if (d <= 0.0) return 0.5 * s;
  return s < d * d ? 0.5 * s : d * sqrt(s) - 0.5 * d * d;
}
__device__ __forceinline__ double dhuber(double s, double d) {  // robust_cost_function.cpp:76-87
  if (d <= 0.0) return 0.5;
  return s < d * d ? 0.5 : 0.5 * d / sqrt(s);
}

__device__ __forceinline__ double edge_penalty(const Window& w, int i) {  // boundary edge: i fixed, i+1 free
  return (i < w.F && i + 1 >= w.F) ? w.boundary_penalty : w.imu_penalty;
}
__device__ __forceinline__ double acc_rw_penalty(const Window& w) {
  return w.acc_rw_penalty >= 0.0 ? w.acc_rw_penalty : w.imu_penalty;
}

// Projection of one observation at state s: false when the point is not in front of the camera (z <= 0,
// imu_sba_fixed_vel.cpp:518). pi = imu_from_w p, pc = cam_from_imu pi.
__device__ bool project(const Window& w, int s, int o, double* pi, double* pc) {
  const double* X = w.pose[s] + kPS * w.okid[o];
  const double* p = w.pts[s] + 3 * w.opid[o];
  const double d[3] = {p[0] - X[kT], p[1] - X[kT + 1], p[2] - X[kT + 2]};
  mtv3(X, d, pi);
  const double* T = w.cam + 12 * w.ocid[o];
  mv3(T, pi, pc);
  pc[0] += T[9]; pc[1] += T[10]; pc[2] += T[11];
  return pc[2] > 0.0;
}

// The inertial residual of edge i (key frames i and i+1) at state s with the bias-corrected deltas
// (GetDeltaRotation / Velocity / Position, imu_preintegration.cpp:162-181; residual imu_sba_fixed_vel.cpp:582-608).
struct EdgeEval {
  double R1T[9], er[3], vt[3], pt[3], r[9], rg[3], ra[3];
};
__device__ void edge_eval(const Window& w, int s, int i, EdgeEval& e) {
  const Preint& pr = w.pre[i];
  const double* A = w.pose[s] + kPS * i;
  const double* B = w.pose[s] + kPS * (i + 1);
  double dbg[3], dba[3], phi[3], E[9], dR[9];
  for (int a = 0; a < 3; ++a) {
    dbg[a] = A[kBg + a] - pr.bg0[a];
    dba[a] = A[kBa + a] - pr.ba0[a];
  }
  mv3(pr.JRg, dbg, phi);
  exp3(phi, E);
  mm3(pr.dR, E, dR);
  polar3(dR);
  double dV[3], dP[3], t1[3], t2[3];
  mv3(pr.JVg, dbg, t1);
  mv3(pr.JVa, dba, t2);
  for (int a = 0; a < 3; ++a) dV[a] = pr.dV[a] + t1[a] + t2[a];
  mv3(pr.JPg, dbg, t1);
  mv3(pr.JPa, dba, t2);
  for (int a = 0; a < 3; ++a) dP[a] = pr.dP[a] + t1[a] + t2[a];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) e.R1T[3 * r + c] = A[3 * c + r];
  double M[9], Q[9];
  mtm3(dR, e.R1T, M);  // dR^T R1^T
  mm3(M, B, Q);        // dR^T R1^T R2
  log3(Q, e.er);
  const double dT = w.pre[i].dT;
  double x[3], y[3];
  for (int a = 0; a < 3; ++a) {
    x[a] = B[kV + a] - A[kV + a] - w.g[a] * dT;
    y[a] = B[kT + a] - A[kT + a] - A[kV + a] * dT - 0.5 * w.g[a] * dT * dT;
  }
  mv3(e.R1T, x, e.vt);
  mv3(e.R1T, y, e.pt);
  for (int a = 0; a < 3; ++a) {
    e.r[a] = e.er[a];
    e.r[3 + a] = e.vt[a] - dV[a];
    e.r[6 + a] = e.pt[a] - dP[a];
    e.rg[a] = A[kBg + a] - B[kBg + a];
    e.ra[a] = A[kBa + a] - B[kBa + a];
  }
}

// d r / d v for variable v in 0..23 ([r1 t1 v1 bg1 ba1 | r2 t2 v2]), row r of the 9-vector residual.
__device__ __forceinline__ double jac(const EdgeModel& m, int r, int v) { return m.J[v / 3][3 * r + v % 3]; }

// Edge i couples key frames i (variables 0..14) and i+1 (15..23).
__device__ __forceinline__ bool edge_active(const Window& w, int i) { return !(i < w.F && i + 1 < w.F); }

}  // namespace

// ---- model ---------------------------------------------------------------------------------------------------------

// Residual, Huber weight and Jacobians of every observation at the current state (imu_sba_fixed_vel.cpp:499-566):
// JR = M [p_imu]x, Jt = -M, Jp = M R^T with M = dproj R_cam_from_imu.
__global__ __launch_bounds__(kThreads) void k_imu_obs_model(Window w) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= w.N) return;
  ObsModel m{};
  double pi[3], pc[3];
  if (project(w, 0, o, pi, pc)) {
    const double iz = 1.0 / pc[2], u = pc[0] * iz, v = pc[1] * iz;
    m.r[0] = u - w.xy[2 * o];
    m.r[1] = v - w.xy[2 * o + 1];
    const double* T = w.cam + 12 * w.ocid[o];
    const double* X = w.pose[0] + kPS * w.okid[o];
    double M[6];
    for (int c = 0; c < 3; ++c) {
      M[c] = iz * T[c] - u * iz * T[6 + c];
      M[3 + c] = iz * T[3 + c] - v * iz * T[6 + c];
    }
    double S[9];
    skew3(pi, S);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) {
        double jr = 0, jp = 0;
        for (int k = 0; k < 3; ++k) {
          jr += M[3 * r + k] * S[3 * k + c];
          jp += M[3 * r + k] * X[3 * c + k];  // (R^T)(k, c) = R(c, k)
        }
        m.JR[3 * r + c] = jr;
        m.Jp[3 * r + c] = jp;
        m.Jt[3 * r + c] = -M[3 * r + c];
      }
    const double* L = w.info + 4 * o;
    const double q = m.r[0] * (L[0] * m.r[0] + L[1] * m.r[1]) + m.r[1] * (L[2] * m.r[0] + L[3] * m.r[1]);
    m.w = dhuber(q, w.scale);
  }
  w.om[o] = m;
}

// Residuals and Jacobians of every inertial edge at the current state (imu_sba_fixed_vel.cpp:575-651), one workgroup
// per edge: its first thread evaluates the edge, then the workgroup forms H = J^T (pen Lambda) J and g = J^T (pen
// Lambda) r, each entry a fixed-order sum.
__device__ void edge_model_eval(const Window& w, int i, EdgeModel& m);

__global__ __launch_bounds__(kEdgeThreads) void k_imu_edge_model(Window w) {
  __shared__ double Js[9][24], LJ[9][25];  // column 24 of LJ: pen Lambda r
  __shared__ double info[81];
  const int i = blockIdx.x, tid = threadIdx.x;
  EdgeModel& m = w.em[i];
  const bool active = edge_active(w, i);
  if (!active) {
    if (tid == 0) m.active = 0;
    return;
  }
  if (tid == 0) edge_model_eval(w, i, m);
  if (tid < 81) info[tid] = w.pre[i].info[tid];
  __syncthreads();  // m (global memory, written by thread 0) is visible to the workgroup
  if (tid < 216) Js[tid / 24][tid % 24] = jac(m, tid / 24, tid % 24);
  __syncthreads();
  const double pen = edge_penalty(w, i);
  if (tid < 225) {
    const int r = tid / 25, v = tid % 25;
    double s = 0.0;
    for (int q = 0; q < 9; ++q) s += info[9 * r + q] * (v < 24 ? Js[q][v] : m.r[q]);
    LJ[r][v] = pen * s;
  }
  __syncthreads();
  for (int e = tid; e < 24 * 25; e += blockDim.x) {
    const int a = e / 25, b = e % 25;
    double s = 0.0;
    for (int r = 0; r < 9; ++r) s += Js[r][a] * LJ[r][b];
    if (b < 24)
      m.H[a][b] = s;
    else
      m.g[a] = s;
  }
}

__device__ void edge_model_eval(const Window& w, int i, EdgeModel& m) {
  m.active = 1;
  EdgeEval e;
  edge_eval(w, 0, i, e);
  for (int a = 0; a < 9; ++a) m.r[a] = e.r[a];
  for (int a = 0; a < 3; ++a) {
    m.rg[a] = e.rg[a];
    m.ra[a] = e.ra[a];
  }
  for (int j = 0; j < 8; ++j)
    for (int a = 0; a < 27; ++a) m.J[j][a] = 0.0;
  const Preint& pr = w.pre[i];
  const double* A = w.pose[0] + kPS * i;
  const double* B = w.pose[0] + kPS * (i + 1);
  auto put = [&](int j, int row0, const double* M33, double s) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m.J[j][3 * (row0 + r) + c] = s * M33[3 * r + c];
  };
  double Jri[9], T1[9], T2[9], S[9];
  jlinv3(e.er, -1.0, Jri);  // J_r^-1(e) = J_l^-1(-e)
  // JR_left rotation rows: -J_r^-1(e) R2^T R1
  double R2T[9], R1[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      R2T[3 * r + c] = B[3 * c + r];
      R1[3 * r + c] = A[3 * r + c];
    }
  mm3(Jri, R2T, T1);
  mm3(T1, R1, T2);
  put(0, 0, T2, -1.0);
  skew3(e.vt, S);
  put(0, 3, S, 1.0);
  skew3(e.pt, S);
  put(0, 6, S, 1.0);
  const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  put(1, 6, I3, -1.0);
  put(2, 3, e.R1T, -1.0);
  put(2, 6, e.R1T, -pr.dT);
  put(4, 3, pr.JVa, -1.0);
  put(4, 6, pr.JPa, -1.0);
  // gyroscope bias, rotation rows: -J_l^-1(e) J_r(JRg dbg) JRg, dbg the stored updated-minus-integration bias
  double phi[3], Jl[9], Jr[9];
  mv3(pr.JRg, pr.dbg, phi);
  jlinv3(e.er, 1.0, Jl);
  jl3(phi, -1.0, Jr);
  mm3(Jl, Jr, T1);
  mm3(T1, pr.JRg, T2);
  put(3, 0, T2, -1.0);
  put(3, 3, pr.JVg, -1.0);
  put(3, 6, pr.JPg, -1.0);
  put(5, 0, Jri, 1.0);
  double R12[9];
  mm3(e.R1T, B, R12);
  put(6, 6, R12, 1.0);
  put(7, 3, e.R1T, 1.0);
}

// ---- full system ---------------------------------------------------------------------------------------------------

// Point blocks V, e_b and the W block of each (point, free key frame) pair, each summed over its observations in
// pobs order (imu_sba_fixed_vel.cpp:701-746).
__global__ __launch_bounds__(kThreads) void k_imu_point_full(Window w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.P) return;
  double V[9] = {}, eb[3] = {};
  for (int q = w.point_obs[i]; q < w.point_obs[i + 1]; ++q) {
    const int o = w.pobs[q];
    const ObsModel& m = w.om[o];
    const double* L = w.info + 4 * o;
    double LJ[2][3], Lr[2];
    for (int r = 0; r < 2; ++r) {
      for (int c = 0; c < 3; ++c) LJ[r][c] = L[2 * r] * m.Jp[c] + L[2 * r + 1] * m.Jp[3 + c];
      Lr[r] = L[2 * r] * m.r[0] + L[2 * r + 1] * m.r[1];
    }
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) V[3 * a + b] += m.w * (m.Jp[a] * LJ[0][b] + m.Jp[3 + a] * LJ[1][b]);
      eb[a] -= m.w * (m.Jp[a] * Lr[0] + m.Jp[3 + a] * Lr[1]);
    }
  }
  for (int a = 0; a < 9; ++a) w.V[9 * i + a] = V[a];
  for (int a = 0; a < 3; ++a) w.eb[3 * i + a] = eb[a];
  for (int p = w.point_pair[i]; p < w.point_pair[i + 1]; ++p) {
    double Wp[18] = {};
    for (int q = w.pair_obs[2 * p]; q < w.pair_obs[2 * p + 1]; ++q) {
      const int o = w.pobs[q];
      const ObsModel& m = w.om[o];
      const double* L = w.info + 4 * o;
      double Jc[2][6];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
          Jc[r][c] = m.JR[3 * r + c];
          Jc[r][3 + c] = m.Jt[3 * r + c];
        }
      double LJ[2][6];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 6; ++c) LJ[r][c] = L[2 * r] * Jc[0][c] + L[2 * r + 1] * Jc[1][c];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 6; ++b) Wp[6 * a + b] += m.w * (m.Jp[a] * LJ[0][b] + m.Jp[3 + a] * LJ[1][b]);
    }
    for (int a = 0; a < 18; ++a) w.W[18 * p + a] = Wp[a];
  }
}

// Row block j (15 rows, all 15n columns) of the pose system U and e_a of free key frame j: its reprojection terms,
// the inertial and random-walk terms of edges k-1 and k (k = F + j), the bias priors of the first free key frame
// (imu_sba_fixed_vel.cpp:701-1003 with the GPU bundler's penalties, sba_imu_v1.cu:1521-1805). Every element is
// written by one thread: no atomics, no separate clearing. The reprojection sums run over the key frame's
// observations in kSlices strided slices, added in slice order; the edge terms are the H and g blocks that
// k_imu_edge_model formed.
__global__ __launch_bounds__(kPoseFullThreads) void k_imu_pose_full(Window w) {
  constexpr int kSlices = kPoseFullThreads / 42;
  __shared__ double part[42 * kSlices];  // the reprojection block (6x6) and right-hand side (6), per slice
  __shared__ double vis[42];
  const int j = blockIdx.x, k = w.F + j, tid = threadIdx.x;
  const int n15 = 15 * w.n;
  if (tid < 42 * kSlices) {  // value tid % 42 over the observations q = slice (mod kSlices)
    const int val = tid % 42, slice = tid / 42;
    const int a = val < 36 ? val / 6 : val - 36, b = val % 6;
    double s = 0.0;
    for (int q = w.pose_obs[j] + slice; q < w.pose_obs[j + 1]; q += kSlices) {
      const int o = w.kobs[q];
      const ObsModel& m = w.om[o];
      const double* L = w.info + 4 * o;
      const double ja0 = a < 3 ? m.JR[a] : m.Jt[a - 3], ja1 = a < 3 ? m.JR[3 + a] : m.Jt[a];
      if (val < 36) {
        const double jb0 = b < 3 ? m.JR[b] : m.Jt[b - 3], jb1 = b < 3 ? m.JR[3 + b] : m.Jt[b];
        s += m.w * (ja0 * (L[0] * jb0 + L[1] * jb1) + ja1 * (L[2] * jb0 + L[3] * jb1));
      } else {
        s -= m.w * (ja0 * (L[0] * m.r[0] + L[1] * m.r[1]) + ja1 * (L[2] * m.r[0] + L[3] * m.r[1]));
      }
    }
    part[tid] = s;
  }
  __syncthreads();
  if (tid < 42) {
    double s = 0.0;
    for (int sl = 0; sl < kSlices; ++sl) s += part[42 * sl + tid];
    vis[tid] = s;
  }
  __syncthreads();
  const bool has_prev = k - 1 >= 0 && edge_active(w, k - 1);
  const bool has_next = k + 1 < w.K;
  const double pen_prev = has_prev ? edge_penalty(w, k - 1) : 0.0, pen_next = has_next ? edge_penalty(w, k) : 0.0;
  const double apen = acc_rw_penalty(w);
  const EdgeModel* ep = has_prev ? w.em + (k - 1) : nullptr;
  const EdgeModel* en = has_next ? w.em + k : nullptr;
  const Preint* pp = has_prev ? w.pre + (k - 1) : nullptr;
  const Preint* pn = has_next ? w.pre + k : nullptr;
  for (int e = tid; e < 15 * n15; e += blockDim.x) {
    const int a = e / n15, c = e % n15, cb = c / 15, b = c % 15;
    double v = 0.0;
    const bool rw = a >= 9 && b >= 9 && (a < 12) == (b < 12);  // same random-walk block (gyroscope or accelerometer)
    const bool gyro = a < 12;
    const int ra = gyro ? a - 9 : a - 12, rb = gyro ? b - 9 : b - 12;
    if (cb == j) {
      if (a < 6 && b < 6) v += vis[6 * a + b];
      if (has_prev && a < 9 && b < 9) v += ep->H[15 + a][15 + b];
      if (has_next) v += en->H[a][b];
      if (rw) {
        if (has_prev) v += gyro ? pen_prev * pp->igrw[3 * ra + rb] : apen * pp->iarw[3 * ra + rb];
        if (has_next) v += gyro ? pen_next * pn->igrw[3 * ra + rb] : apen * pn->iarw[3 * ra + rb];
      }
      if (j == 0 && a == b && a >= 9) v += a < 12 ? w.prior_gyro : w.prior_acc;
    } else if (cb == j - 1 && has_prev) {  // key frame k-1 is free: cross terms of edge k-1 (k on its right)
      if (a < 9) v += ep->H[15 + a][b];
      if (rw) v -= gyro ? pen_prev * pp->igrw[3 * ra + rb] : apen * pp->iarw[3 * ra + rb];
    } else if (cb == j + 1 && has_next) {  // cross terms of edge k (k on its left)
      if (b < 9) v += en->H[a][15 + b];
      if (rw) v -= gyro ? pen_next * pn->igrw[3 * ra + rb] : apen * pn->iarw[3 * ra + rb];
    }
    w.U[static_cast<int64_t>(15 * j + a) * n15 + c] = v;
  }
  if (tid < 15) {
    const int a = tid;
    double v = a < 6 ? vis[36 + a] : 0.0;
    const double* X = w.pose[0] + kPS * k;
    if (has_prev && a < 9) v -= ep->g[15 + a];
    if (has_next) v -= en->g[a];
    if (a >= 9) {
      const bool gyro = a < 12;
      const int ra = gyro ? a - 9 : a - 12;
      if (has_prev) {  // k is the right key frame of edge k-1: d r_rw / d b_k = -I
        const double* I3 = gyro ? pp->igrw : pp->iarw;
        const double* r = gyro ? ep->rg : ep->ra;
        v += (gyro ? pen_prev : apen) * (I3[3 * ra] * r[0] + I3[3 * ra + 1] * r[1] + I3[3 * ra + 2] * r[2]);
      }
      if (has_next) {
        const double* I3 = gyro ? pn->igrw : pn->iarw;
        const double* r = gyro ? en->rg : en->ra;
        v -= (gyro ? pen_next : apen) * (I3[3 * ra] * r[0] + I3[3 * ra + 1] * r[1] + I3[3 * ra + 2] * r[2]);
      }
      if (j == 0) v -= gyro ? w.prior_gyro * X[kBg + ra] : w.prior_acc * X[kBa + ra];
    }
    w.ea[15 * j + a] = v;
  }
}

// ---- reduced system ------------------------------------------------------------------------------------------------

// The reference's pivoted LDLT inverse of a symmetric 3x3 block (cuda_sba_common.h:365-470): the largest diagonal
// first, then the larger remaining one (compared before its update); negative pivots are clipped to zero and a zero
// pivot contributes nothing, so a point without observations steps zero.
__device__ void ldlt_inverse3(const double* A, double* X) {
  double a00 = A[0], a11 = A[4], a22 = A[8], a10 = A[3], a20 = A[6], a21 = A[7];
  bool p01 = false, p02 = false, p12 = false;
  const double d0 = fabs(a00), d1 = fabs(a11), d2 = fabs(a22);
  double t;
  if (fmax(d1, d2) > d0) {
    if (d1 > d2) {
      t = a20; a20 = a21; a21 = t;
      t = a00; a00 = a11; a11 = t;
      p01 = true;
    } else {
      t = a00; a00 = a22; a22 = t;
      t = a10; a10 = a21; a21 = t;
      p02 = true;
    }
  }
  a00 = fmax(a00, 0.0);
  a10 = a00 == 0.0 ? 0.0 : a10 / a00;
  a20 = a00 == 0.0 ? 0.0 : a20 / a00;
  if (fabs(a22) > fabs(a11)) {
    t = a10; a10 = a20; a20 = t;
    t = a11; a11 = a22; a22 = t;
    p12 = true;
  }
  a11 = fmax(a11 - a10 * a10 * a00, 0.0);
  a21 = a11 == 0.0 ? 0.0 : (a21 - a20 * a10 * a00) / a11;
  a22 = fmax(a22 - a20 * a20 * a00 - a21 * a21 * a11, 0.0);
  const double i0 = a00 == 0.0 ? 0.0 : 1.0 / a00, i1 = a11 == 0.0 ? 0.0 : 1.0 / a11;
  const double i2 = a22 == 0.0 ? 0.0 : 1.0 / a22;
  for (int c = 0; c < 3; ++c) {
    double b[3] = {c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, c == 2 ? 1.0 : 0.0};
    if (p01) { t = b[0]; b[0] = b[1]; b[1] = t; }
    if (p02) { t = b[0]; b[0] = b[2]; b[2] = t; }
    if (p12) { t = b[1]; b[1] = b[2]; b[2] = t; }
    const double y0 = b[0], y1 = b[1] - y0 * a10, y2 = b[2] - y0 * a20 - y1 * a21;
    double x[3];
    x[2] = y2 * i2;
    x[1] = y1 * i1 - x[2] * a21;
    x[0] = y0 * i0 - x[1] * a10 - x[2] * a20;
    if (p12) { t = x[1]; x[1] = x[2]; x[2] = t; }
    if (p02) { t = x[0]; x[0] = x[2]; x[2] = t; }
    if (p01) { t = x[0]; x[0] = x[1]; x[1] = t; }
    X[c] = x[0]; X[3 + c] = x[1]; X[6 + c] = x[2];
  }
}

// V* = V + lambda diag V inverted (sba_imu_v1.cu:247-294), Vinv e_b, and Y = Vinv W for each pair of the point.
__global__ __launch_bounds__(kThreads) void k_imu_point_inv(Window w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.P) return;
  double Vd[9], X[9];
  for (int a = 0; a < 9; ++a) Vd[a] = w.V[9 * i + a];
  for (int a = 0; a < 3; ++a) Vd[4 * a] += w.lambda * w.V[9 * i + 4 * a];
  ldlt_inverse3(Vd, X);
  for (int a = 0; a < 9; ++a) w.Vinv[9 * i + a] = X[a];
  double e[3] = {w.eb[3 * i], w.eb[3 * i + 1], w.eb[3 * i + 2]};
  for (int a = 0; a < 3; ++a) w.rp[3 * i + a] = X[3 * a] * e[0] + X[3 * a + 1] * e[1] + X[3 * a + 2] * e[2];
  for (int p = w.point_pair[i]; p < w.point_pair[i + 1]; ++p) {
    const double* Wp = w.W + 18 * p;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 6; ++b)
        w.Y[18 * p + 6 * a + b] = X[3 * a] * Wp[b] + X[3 * a + 1] * Wp[6 + b] + X[3 * a + 2] * Wp[12 + b];
  }
}

// Block (j, k), j >= k, of S = U + lambda diag U - sum_i W_ij^T Vinv_i W_ik, one workgroup per block; workgroups
// past the blocks form the reduced right-hand side e_a - sum_i W_ij^T Vinv_i e_b,i of one key frame each. The sums
// over points are split over kReducedSlices strided slices (slice s takes every kReducedSlices-th term from the s-th)
// and the slices are then added in order, so every entry is a fixed-order sum.
__global__ __launch_bounds__(kReducedThreads) void k_imu_reduced(Window w) {
  __shared__ double part[36 * kReducedSlices];
  const int n15 = 15 * w.n, tid = threadIdx.x;
  if (blockIdx.x < w.nblocks) {  // workgroup-uniform
    const int j = w.blk_jk[2 * blockIdx.x], k = w.blk_jk[2 * blockIdx.x + 1];
    if (tid < 36 * kReducedSlices) {
      const int val = tid % 36, slice = tid / 36, a = val / 6, c = val % 6;
      double s = 0.0;
      for (int t = w.blk_tri[blockIdx.x] + slice; t < w.blk_tri[blockIdx.x + 1]; t += kReducedSlices) {
        const int2 pr = w.tri[t];
        const double* Wj = w.W + 18 * pr.x;
        const double* Yk = w.Y + 18 * pr.y;
        s += Wj[a] * Yk[c] + Wj[6 + a] * Yk[6 + c] + Wj[12 + a] * Yk[12 + c];
      }
      part[tid] = s;
    }
    __syncthreads();
    if (tid >= 256) return;
    const int a = tid / 16, c = tid % 16;
    double* T = w.A + 256 * static_cast<int64_t>(blockIdx.x);  // lower tile (j, k), blocks in the same order
    if (a == 15 || c == 15) {  // padding: zero, and one on the diagonal of a diagonal tile
      T[tid] = (a == c && j == k) ? 1.0 : 0.0;
      return;
    }
    const int64_t row = 15 * j + a, col = 15 * k + c;
    double v = w.U[row * n15 + col];
    if (row == col) v += w.lambda * v;
    if (a < 6 && c < 6) {
      double s = 0.0;
      for (int sl = 0; sl < kReducedSlices; ++sl) s += part[36 * sl + 6 * a + c];
      v -= s;
    }
    T[tid] = v;
    return;
  }
  const int j = blockIdx.x - w.nblocks;
  if (tid < 6 * kReducedSlices) {
    const int val = tid % 6, slice = tid / 6;
    double s = 0.0;
    for (int q = w.pose_pair[j] + slice; q < w.pose_pair[j + 1]; q += kReducedSlices) {
      const int p = w.kpair[q];
      const double* Wp = w.W + 18 * p;
      const double* r = w.rp + 3 * w.pair_point[p];
      s += Wp[val] * r[0] + Wp[6 + val] * r[1] + Wp[12 + val] * r[2];
    }
    part[tid] = s;
  }
  __syncthreads();
  if (tid >= 15) return;
  double v = w.ea[15 * j + tid];
  if (tid < 6) {
    double s = 0.0;
    for (int sl = 0; sl < kReducedSlices; ++sl) s += part[6 * sl + tid];
    v -= s;
  }
  w.rhs[15 * j + tid] = v;
}

// Block Cholesky S = L L^T of the reduced system (right-looking; Golub & Van Loan, Matrix Computations, Algorithm
// 4.2.10), over 16 x 16 FP64 tiles that hold the 15 x 15 blocks of the free key frames padded with a unit diagonal
// (lower tile (i, j) at A + 256 (i (i + 1) / 2 + j), row-major, i.e. in the matrix-core register layout). It keeps the
// tile sparsity pattern of L: the inertial terms couple neighbouring key frames and the reprojections the key frames
// that see common points, so most tiles stay zero and are skipped (fill-in is tracked). Per tile column k:
//   1. wavefront 0 factors the diagonal tile and forms X = L_kk^-1 (sba_tiles::factor_tile), stores X^T over the
//      diagonal tile for the backward substitution, and steps the forward substitution: y_k = X b_k;
//   2. the panel tiles below it, one wavefront each: L_ik = A_ik X^T on the matrix cores, kept transposed in LDS,
//      and b_i -= L_ik y_k;
//   3. the trailing tiles A_ab -= L_ak L_bk^T of every pair of panel rows, one wavefront per tile, on the matrix cores.
// Then the backward substitution x_k = X_k^T (y_k - sum_{i > k} L_ik^T x_i), from the last tile column (column
// sweeps: x_k is formed, then y_i -= L_ki^T x_k for the tiles left of it). A non-positive pivot clears chol_ok (the LM
// loop raises lambda, imu_sba_gpu_fixed_vel.cpp:199-203). The step overwrites rhs. One workgroup; every barrier is
// reached by every thread (loop bounds and the pivot flag are workgroup-uniform, read from LDS after a barrier).
__global__ __launch_bounds__(kFactorThreads) void k_imu_factor(Window w) {
  using namespace sba_tiles;
  __shared__ double PT[(kMaxFree - 1) * 256];         // panel tiles of the current column, transposed
  __shared__ double sT[256], xt[256], yk[16], rdiag[16];
  __shared__ double x[16 * kMaxFree];                 // right-hand side, then y, then the step (padded by tiles)
  __shared__ unsigned char nz[kMaxFree * kMaxFree];  // nz[i * nb + j], i >= j: tile (i, j) of L is not zero
  __shared__ int rows[kMaxFree];
  __shared__ int nrows, ok;
  const int nb = w.n, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  auto tile = [&](int i, int j) { return w.A + 256 * (static_cast<int64_t>(i) * (i + 1) / 2 + j); };
  auto decode = [](int b, int& i, int& j) {  // b = i (i + 1) / 2 + j, j <= i
    i = static_cast<int>((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > b) --i;
    while ((i + 1) * (i + 2) / 2 <= b) ++i;
    j = b - i * (i + 1) / 2;
  };
  for (int b = wave; b < nb * (nb + 1) / 2; b += nwaves) {
    int i, j;
    decode(b, i, j);
    const double* T = tile(i, j);
    bool any = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) any |= T[64 * v + lane] != 0.0;
    const bool blk = __any(any);
    if (lane == 0) nz[i * nb + j] = (blk || i == j) ? 1 : 0;
  }
  for (int i = tid; i < 16 * nb; i += blockDim.x) x[i] = (i & 15) < 15 ? w.rhs[15 * (i >> 4) + (i & 15)] : 0.0;
  if (tid == 0) ok = 1;
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    if (wave == 0) {
      double* D = tile(k, k);
      double4v t;
#pragma unroll
      for (int v = 0; v < 4; ++v) t[v] = D[64 * v + lane];
      const bool pos = factor_tile(t, sT, xt, rdiag, lane);
      wave_lds_sync();
#pragma unroll
      for (int v = 0; v < 4; ++v) D[64 * v + lane] = xt[64 * v + lane];
      if (lane < 16) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 16; ++m) s += xt[16 * m + lane] * x[16 * k + m];  // X[lane][m] = X^T[m][lane]
        yk[lane] = s;
      }
      wave_lds_sync();
      if (lane < 16) x[16 * k + lane] = yk[lane];
      if (lane == 0) {
        if (!pos) ok = 0;
        int n = 0;
        for (int i = k + 1; i < nb; ++i)
          if (nz[i * nb + k]) rows[n++] = i;
        nrows = n;
      }
    }
    __syncthreads();
    if (!ok) break;
    for (int p = wave; p < nrows; p += nwaves) {
      double* T = tile(rows[p], k);
      double4v c = {0.0, 0.0, 0.0, 0.0};
      // A operand: A_ik[l % 16][4 q + l / 16]; B operand: X^T[4 q + l / 16][l % 16]
#pragma unroll
      for (int q = 0; q < 4; ++q) c = mfma64(T[16 * (lane & 15) + 4 * q + (lane >> 4)], xt[64 * q + lane], c);
      double* P = PT + 256 * p;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        T[64 * v + lane] = c[v];
        P[16 * (lane & 15) + 4 * v + (lane >> 4)] = c[v];  // element (4 v + l / 16, l % 16) transposed
      }
      wave_lds_sync();
      if (lane < 16) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 16; ++m) s += P[16 * m + lane] * yk[m];  // L_ik[lane][m]
        x[16 * rows[p] + lane] -= s;
      }
    }
    __syncthreads();
    const int npairs = nrows * (nrows + 1) / 2;
    for (int q = wave; q < npairs; q += nwaves) {
      int a, b;
      decode(q, a, b);
      const int ia = rows[a], ib = rows[b];
      double* C = tile(ia, ib);
      double4v c;
#pragma unroll
      for (int v = 0; v < 4; ++v) c[v] = C[64 * v + lane];
      c = tile_update(c, PT + 256 * a, PT + 256 * b, lane);  // C - L_ak L_bk^T
#pragma unroll
      for (int v = 0; v < 4; ++v) C[64 * v + lane] = c[v];
      if (lane == 0) nz[ia * nb + ib] = 1;
    }
    __syncthreads();
  }
  if (!ok) {
    if (tid == 0) *w.chol_ok = 0;
    return;
  }
  for (int k = nb - 1; k >= 0; --k) {
    if (wave == 0) {
      if (lane < 16) {
        const double* XT = tile(k, k);
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 16; ++m) s += XT[16 * lane + m] * x[16 * k + m];
        yk[lane] = s;
      }
      wave_lds_sync();
      if (lane < 16) x[16 * k + lane] = yk[lane];
    }
    __syncthreads();
    for (int i = wave; i < k; i += nwaves) {
      if (!nz[k * nb + i] || lane >= 16) continue;
      const double* T = tile(k, i);
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 16; ++m) s += T[16 * m + lane] * x[16 * k + m];  // L_ki[m][lane]
      x[16 * i + lane] -= s;
    }
    __syncthreads();
  }
  for (int i = tid; i < 16 * nb; i += blockDim.x)
    if ((i & 15) < 15) w.rhs[15 * (i >> 4) + (i & 15)] = x[i];
  if (tid == 0) *w.chol_ok = 1;
}

// dp = Vinv (e_b - sum_j W_ij dc_j), the trial point, and the point's share of the prediction
// dp^T (V + 2 lambda diag V) dp + 2 dp^T sum_j W_ij dc_j (sba_imu_v1.cu:372-394, 767-819).
__global__ __launch_bounds__(kThreads) void k_imu_backsub(Window w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.P || !*w.chol_ok) return;
  double t[3] = {0, 0, 0};
  for (int p = w.point_pair[i]; p < w.point_pair[i + 1]; ++p) {
    const double* Wp = w.W + 18 * p;
    const double* dc = w.rhs + 15 * w.pair_pose[p];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 6; ++b) t[a] += Wp[6 * a + b] * dc[b];
  }
  const double* X = w.Vinv + 9 * i;
  double e[3], dp[3];
  for (int a = 0; a < 3; ++a) e[a] = w.eb[3 * i + a] - t[a];
  double mx = 0.0;
  for (int a = 0; a < 3; ++a) {
    dp[a] = X[3 * a] * e[0] + X[3 * a + 1] * e[1] + X[3 * a + 2] * e[2];
    w.dp[3 * i + a] = dp[a];
    w.pts[1][3 * i + a] = w.pts[0][3 * i + a] + dp[a];
    mx = fmax(mx, fabs(dp[a]));
  }
  const double* V = w.V + 9 * i;
  const double l2 = 1.0 + 2.0 * w.lambda;
  double pred = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) pred += dp[a] * V[3 * a + b] * dp[b] * (a == b ? l2 : 1.0);
  pred += 2.0 * (dp[0] * t[0] + dp[1] * t[1] + dp[2] * t[2]);
  w.ppred[i] = pred;
  w.pmax[i] = mx;
}

// Trial key frames: R Exp(dr), t + R dt, v + dv, biases + db (imu_sba_fixed_vel.cpp:224-240).
__global__ __launch_bounds__(64) void k_imu_pose_trial(Window w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w.n || !*w.chol_ok) return;
  const int k = w.F + j;
  const double* d = w.rhs + 15 * j;
  const double* X = w.pose[0] + kPS * k;
  double* Y = w.pose[1] + kPS * k;
  double E[9], R[9], dt[3];
  exp3(d, E);
  mm3(X, E, R);
  mv3(X, d + 3, dt);
  for (int a = 0; a < 9; ++a) Y[a] = R[a];
  for (int a = 0; a < 3; ++a) {
    Y[kT + a] = X[kT + a] + dt[a];
    Y[kV + a] = X[kV + a] + d[6 + a];
    Y[kBg + a] = X[kBg + a] + d[9 + a];
    Y[kBa + a] = X[kBa + a] + d[12 + a];
  }
}

// Row r of dc^T (U + 2 lambda diag U) dc (sba_imu_v1.cu:821-852), one wavefront per row. U couples a key frame only
// with its neighbours (the inertial edges), so row r is summed over the column blocks of key frames j-1..j+1, which
// hold every non-zero of the row (at most 45 columns, one per lane); the lanes are added by a fixed shuffle tree.
__global__ __launch_bounds__(kThreads) void k_imu_pose_pred(Window w) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
  const int M = 15 * w.n;
  if (r >= M || !*w.chol_ok) return;  // wavefront-uniform
  const int jb = r / 15;
  const int c0 = jb > 0 ? 15 * (jb - 1) : 0, c1 = min(M, 15 * (jb + 2));
  const int c = c0 + lane;
  const double* dc = w.rhs;
  double s = 0.0;
  if (c < c1) s = w.U[static_cast<int64_t>(r) * M + c] * dc[c] * (c == r ? 1.0 + 2.0 * w.lambda : 1.0);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) w.cpred[r] = dc[r] * s;
}

// ---- cost ----------------------------------------------------------------------------------------------------------

// The cost terms of state s (sba_imu_v1.cu:416-748): thread o < N the Huber term of observation o (-1 when it is
// skipped), thread N + i the inertial and random-walk terms of edge i, thread N + K - 1 the bias priors.
__global__ __launch_bounds__(kThreads) void k_imu_cost(Window w, int s) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 1 && !*w.chol_ok) return;
  if (t < w.N) {
    double pi[3], pc[3];
    if (!project(w, s, t, pi, pc)) {
      w.ocost[t] = -1.0;
      return;
    }
    const double r0 = pc[0] / pc[2] - w.xy[2 * t], r1 = pc[1] / pc[2] - w.xy[2 * t + 1];
    const double* L = w.info + 4 * t;
    w.ocost[t] = huber(r0 * (L[0] * r0 + L[1] * r1) + r1 * (L[2] * r0 + L[3] * r1), w.scale);
    return;
  }
  const int i = t - w.N;
  if (i < w.K - 1) {
    if (!edge_active(w, i)) {
      w.ecost[i] = 0.0;
      return;
    }
    EdgeEval e;
    edge_eval(w, s, i, e);
    const Preint& pr = w.pre[i];
    double q = 0.0, qg = 0.0, qa = 0.0;
    for (int a = 0; a < 9; ++a) {
      double u = 0.0;
      for (int b = 0; b < 9; ++b) u += pr.info[9 * a + b] * e.r[b];
      q += e.r[a] * u;
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        qg += e.rg[a] * pr.igrw[3 * a + b] * e.rg[b];
        qa += e.ra[a] * pr.iarw[3 * a + b] * e.ra[b];
      }
    const double pen = edge_penalty(w, i);
    w.ecost[i] = pen * qg + acc_rw_penalty(w) * qa + pen * q;
  } else if (i == w.K - 1) {
    double c = 0.0;
    if (w.F < w.K) {
      const double* X = w.pose[s] + kPS * w.F;
      for (int a = 0; a < 3; ++a)
        c += w.prior_gyro * X[kBg + a] * X[kBg + a] + w.prior_acc * X[kBa + a] * X[kBa + a];
    }
    w.ecost[w.K - 1] = c;
  }
}

// The iteration's scalars in one workgroup, in a fixed order: each thread sums a fixed stride of every stream, then
// the threads' sums combine in a fixed tree.
__global__ __launch_bounds__(1024) void k_imu_reduce(Window w, int with_step) {
  __shared__ double sc[1024], sp[1024], smp[1024], smc[1024];
  __shared__ int sk[1024];
  const int tid = threadIdx.x;
  const bool step = with_step && *w.chol_ok;
  double c = 0.0, pr = 0.0, mp = 0.0, mc = 0.0;
  int skipped = 0;
  if (!with_step || step) {
    for (int o = tid; o < w.N; o += 1024) {
      const double v = w.ocost[o];
      if (v < 0.0) ++skipped; else c += v;
    }
    for (int i = tid; i < w.K; i += 1024) c += w.ecost[i];
  }
  if (step) {
    for (int i = tid; i < w.P; i += 1024) {
      pr += w.ppred[i];
      mp = fmax(mp, w.pmax[i]);
    }
    for (int r = tid; r < 15 * w.n; r += 1024) {
      pr += w.cpred[r];
      mc = fmax(mc, fabs(w.rhs[r]));
    }
  }
  sc[tid] = c; sp[tid] = pr; smp[tid] = mp; smc[tid] = mc; sk[tid] = skipped;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (tid < h) {
      sc[tid] += sc[tid + h];
      sp[tid] += sp[tid + h];
      smp[tid] = fmax(smp[tid], smp[tid + h]);
      smc[tid] = fmax(smc[tid], smc[tid + h]);
      sk[tid] += sk[tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    IterScalars o{};
    o.cost = sc[0];
    o.pred = sp[0];
    o.maxdp = smp[0];
    o.maxdc = smc[0];
    o.skipped = sk[0];
    o.chol_ok = with_step ? *w.chol_ok : 1;
    *w.out = o;
  }
}

// The trial state becomes the current one.
__global__ __launch_bounds__(kThreads) void k_imu_accept(Window w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * w.P) w.pts[0][t] = w.pts[1][t];
  if (t < kPS * w.K) w.pose[0][t] = w.pose[1][t];
}

namespace sba_imu_dev {

static inline unsigned blocks(int n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

hipError_t imu_model(const Window& w, hipStream_t s) {
  if (w.N > 0) hipLaunchKernelGGL(k_imu_obs_model, dim3(blocks(w.N, kThreads)), dim3(kThreads), 0, s, w);
  if (w.K > 1) hipLaunchKernelGGL(k_imu_edge_model, dim3(w.K - 1), dim3(kEdgeThreads), 0, s, w);
  if (w.P > 0) hipLaunchKernelGGL(k_imu_point_full, dim3(blocks(w.P, kThreads)), dim3(kThreads), 0, s, w);
  if (w.n > 0) hipLaunchKernelGGL(k_imu_pose_full, dim3(w.n), dim3(kPoseFullThreads), 0, s, w);
  return hipGetLastError();
}

hipError_t imu_step(const Window& w, hipStream_t s) {
  if (w.P > 0) hipLaunchKernelGGL(k_imu_point_inv, dim3(blocks(w.P, kThreads)), dim3(kThreads), 0, s, w);
  if (w.n > 0) hipLaunchKernelGGL(k_imu_reduced, dim3(w.nblocks + w.n), dim3(kReducedThreads), 0, s, w);
  hipLaunchKernelGGL(k_imu_factor, dim3(1), dim3(kFactorThreads), 0, s, w);  // also sets chol_ok when n = 0
  if (w.P > 0) hipLaunchKernelGGL(k_imu_backsub, dim3(blocks(w.P, kThreads)), dim3(kThreads), 0, s, w);
  if (w.n > 0) {
    hipLaunchKernelGGL(k_imu_pose_trial, dim3(blocks(w.n, 64)), dim3(64), 0, s, w);
    hipLaunchKernelGGL(k_imu_pose_pred, dim3(blocks(15 * w.n, kThreads / 64)), dim3(kThreads), 0, s, w);
  }
  hipLaunchKernelGGL(k_imu_cost, dim3(blocks(w.N + w.K, kThreads)), dim3(kThreads), 0, s, w, 1);
  hipLaunchKernelGGL(k_imu_reduce, dim3(1), dim3(1024), 0, s, w, 1);
  return hipGetLastError();
}

hipError_t imu_cost(const Window& w, hipStream_t s) {
  hipLaunchKernelGGL(k_imu_cost, dim3(blocks(w.N + w.K, kThreads)), dim3(kThreads), 0, s, w, 0);
  hipLaunchKernelGGL(k_imu_reduce, dim3(1), dim3(1024), 0, s, w, 0);
  return hipGetLastError();
}

hipError_t imu_accept(const Window& w, hipStream_t s) {
  const int n = 3 * w.P > kPS * w.K ? 3 * w.P : kPS * w.K;
  hipLaunchKernelGGL(k_imu_accept, dim3(blocks(n, kThreads)), dim3(kThreads), 0, s, w);
  return hipGetLastError();
}

}  // namespace sba_imu_dev
