// TEST INFRASTRUCTURE -- the CPU oracle of the sliding-window SBA. Only tests/, the smoke check and the
// benchmark's CPU-baseline leg load this library, and only as the checker or the timed CPU baseline; the
// product (cuvslam_amd/) never calls it.
//
// A restatement, in plain C++ with no Eigen, of cuVSLAM's Schur-complement Levenberg-Marquardt bundler
// (libs/sba/schur_complement_bundler_cpu.cpp, with the GPU bundler's control flow of
// libs/sba/schur_complement_bundler_gpu.cpp:115-259). It is templated on the scalar:
//   * double: the FP64 oracle of the plain method (Cholesky RCS solve, block-sparse Schur complement);
//   * float : the reference CPU path restated step for step (dense Y*W product as cpu.cpp:264-277, Jacobi
//             pseudo-inverse of the RCS with threshold 1e-6 as cpu.cpp:337-346, FP32 everywhere).
// Each function cites the reference lines it follows. What pins each function (closed forms, finite
// differences, numpy dense solves, the reference's own known answers) is in tests/test_oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <tuple>
#include <vector>

namespace {

constexpr float kFltEps = std::numeric_limits<float>::epsilon();  // common/types.h:203-206
const float kSqrtFltEps = std::sqrt(kFltEps);                       // common/types.h:208-211
constexpr float kHither = 0.1f;                                     // epipolar/near_plane.h:33-38

struct OrcProblem {  // same field order as sba_problem_view (include/cuvslam_sba.h)
  int32_t num_points, num_fixed_points, num_observations, num_poses, num_fixed_key_frames, num_cameras;
  const float* points;
  const float* observation_xys;
  const float* observation_infos;
  const int32_t* point_ids;
  const int32_t* pose_ids;
  const int8_t* camera_ids;
  const float* rig_from_world;
  const float* camera_from_rig;
  int32_t max_iterations;
  float robustifier_scale;
};

struct OrcOptions {
  int32_t loss;              // 0 Student-t, 1 Huber
  int32_t rcs_solver;        // 0 Cholesky (GPU bundler), 1 Jacobi pseudo-inverse (CPU bundler)
  int32_t dense_schur;       // 1: dense temp * point_pose_block product as the CPU bundler does
  int32_t fixed_iterations;  // 1: no step-size convergence exit
  int32_t reorthonormalize;  // 1: UpdateState re-orthonormalizes rotations (cpu.cpp:377-380)
  // Rounding model of an FP32 reduced system (FP64 solver only; a sensitivity probe, not a reference behaviour):
  //   0 none; 1 the assembled S and e rounded to FP32; 2 the camera blocks U, e_a and the Schur sums
  //   sum W V*^-1 W^T, sum W V*^-1 e_b each rounded to FP32 before S = U - sum W V*^-1 W^T (+ damping) is formed
  //   in FP64, the order in which the MI355X kernels combine their FP32 accumulators.
  int32_t round_reduced_f32;
  // Branch probe (FP64 solver only): at this LM iteration (1-based; 0 = never) the damping decision takes the
  // other side of the threshold its gain ratio is nearest to -- accept/reject when rho is nearer 0.25, keep/shrink
  // lambda when it is nearer 0.75. Gives the FP64 trajectory that a solver whose rho lands on the other side of a
  // threshold (through rounding) follows.
  int32_t flip_iteration;
  int32_t pad;
};

struct OrcResult {
  double* points;  // out [np][3]
  double* poses;   // out [nc][16] column-major
  int32_t iterations, status, accepted, pad;
  double initial_cost, last_cost, lambda;
  int32_t trace_capacity, pad2;
  double* trace_lambda;
  double* trace_cost;
  double* trace_rho;
  int32_t* trace_flags;  // bit 0 accepted, bit 1 factorization failed, bit 2 converged exit
};

template <class T>
struct Pose {
  T R[3][3];
  T t[3];
};

template <class T>
Pose<T> compose(const Pose<T>& a, const Pose<T>& b) {  // a * b
  Pose<T> c;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) c.R[i][j] = a.R[i][0] * b.R[0][j] + a.R[i][1] * b.R[1][j] + a.R[i][2] * b.R[2][j];
    c.t[i] = a.R[i][0] * b.t[0] + a.R[i][1] * b.t[1] + a.R[i][2] * b.t[2] + a.t[i];
  }
  return c;
}

template <class T>
void apply(const Pose<T>& a, const T p[3], T q[3]) {
  for (int i = 0; i < 3; ++i) q[i] = a.R[i][0] * p[0] + a.R[i][1] * p[1] + a.R[i][2] * p[2] + a.t[i];
}

template <class T>
Pose<T> load_pose(const float* m) {  // column-major 4x4 (Eigen Isometry3f)
  Pose<T> p;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) p.R[r][c] = static_cast<T>(m[c * 4 + r]);
    p.t[r] = static_cast<T>(m[12 + r]);
  }
  return p;
}

template <class T>
void store_pose(const Pose<T>& p, double* m) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[c * 4 + r] = static_cast<double>(p.R[r][c]);
  m[3] = m[7] = m[11] = 0;
  for (int r = 0; r < 3; ++r) m[12 + r] = static_cast<double>(p.t[r]);
  m[15] = 1;
}

// Robust losses, math/robust_cost_function.cpp:62-95 (Student-t with nu = 1, robust_cost_function.h:34-35).
template <class T>
T loss_rho(int kind, T s, T d) {
  if (kind == 0) return (d + T(1)) * std::log(T(1) + s / d);
  if (d <= T(0)) return T(0.5) * s;
  if (s < d * d) return T(0.5) * s;
  return d * std::sqrt(s) - T(0.5) * d * d;
}

template <class T>
T loss_drho(int kind, T s, T d) {
  if (kind == 0) return (d + T(1)) / (d + s);
  if (d <= T(0)) return T(0.5);
  if (s < d * d) return T(0.5);
  return T(0.5) * d / std::sqrt(s);
}

// Exp of a twist (omega, v), math/twist.cpp:23-41 (theta clamped at FLT_EPSILON).
template <class T>
Pose<T> twist_exp(const T tw[6]) {
  const T* w = tw;
  const T* v = tw + 3;
  const T theta = std::max<T>(std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]), T(kFltEps));
  const T c = std::cos(theta);
  const T st = std::sin(theta) / theta;
  const T n[3] = {w[0] / theta, w[1] / theta, w[2] / theta};
  Pose<T> E;
  const T ws[3] = {w[0] * st, w[1] * st, w[2] * st};
  // SkewSymmetric(w * s_theta, c) + (1 - c) n n^T   (common/types.h:189-191)
  const T K[3][3] = {{c, -ws[2], ws[1]}, {ws[2], c, -ws[0]}, {-ws[1], ws[0], c}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) E.R[i][j] = K[i][j] + (T(1) - c) * n[i] * n[j];
  const T nxv[3] = {n[1] * v[2] - n[2] * v[1], n[2] * v[0] - n[0] * v[2], n[0] * v[1] - n[1] * v[0]};
  const T nxnxv[3] = {n[1] * nxv[2] - n[2] * nxv[1], n[2] * nxv[0] - n[0] * nxv[2], n[0] * nxv[1] - n[1] * nxv[0]};
  for (int i = 0; i < 3; ++i) E.t[i] = v[i] + ((T(1) - c) / theta * nxv[i] + (T(1) - st) * nxnxv[i]);
  return E;
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (Golub & Van Loan, Matrix Computations,
// Algorithm 8.5.3). a (row-major, destroyed) -> eigenvalues in ev, eigenvectors as columns of q.
template <class T>
void sym_jacobi(int n, T* a, T* ev, T* q) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) q[i * n + j] = (i == j);
  T fro = 0;
  for (int i = 0; i < n * n; ++i) fro += a[i] * a[i];
  const T tol = std::numeric_limits<T>::epsilon() * std::sqrt(fro);
  for (int sweep = 0; sweep < 60; ++sweep) {
    T off = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
    if (std::sqrt(T(2) * off) <= tol) break;
    for (int p = 0; p < n; ++p)
      for (int r = p + 1; r < n; ++r) {
        const T apr = a[p * n + r];
        if (apr == T(0)) continue;
        const T tau = (a[r * n + r] - a[p * n + p]) / (T(2) * apr);
        const T t = (tau >= 0 ? T(1) : T(-1)) / (std::fabs(tau) + std::sqrt(T(1) + tau * tau));
        const T c = T(1) / std::sqrt(T(1) + t * t);
        const T s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J (columns p, r)
          const T akp = a[k * n + p], akr = a[k * n + r];
          a[k * n + p] = c * akp - s * akr;
          a[k * n + r] = s * akp + c * akr;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, r)
          const T apk = a[p * n + k], ark = a[r * n + k];
          a[p * n + k] = c * apk - s * ark;
          a[r * n + k] = s * apk + c * ark;
        }
        for (int k = 0; k < n; ++k) {
          const T qkp = q[k * n + p], qkr = q[k * n + r];
          q[k * n + p] = c * qkp - s * qkr;
          q[k * n + r] = s * qkp + c * qkr;
        }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = a[i * n + i];
}

// Pseudo-inverse of a symmetric matrix through its SVD, the way Eigen::JacobiSVD::solve with setThreshold
// works (cpu.cpp:226-254, 339-346): a singular value |lambda_k| counts when it exceeds thr * max |lambda|.
// For a symmetric matrix the singular values are |lambda_k| and the pseudo-inverse is sum q_k q_k^T / lambda_k.
template <class T>
int sym_pinv(int n, const T* A, T thr, T* out) {
  std::vector<T> a(A, A + n * n), ev(n), q(n * n);
  sym_jacobi(n, a.data(), ev.data(), q.data());
  T smax = 0;
  for (int k = 0; k < n; ++k) smax = std::max(smax, std::fabs(ev[k]));
  int rank = 0;
  std::fill(out, out + n * n, T(0));
  if (smax <= T(0)) return 0;
  for (int k = 0; k < n; ++k) {
    if (!(std::fabs(ev[k]) > thr * smax)) continue;
    ++rank;
    const T inv = T(1) / ev[k];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) out[i * n + j] += q[i * n + k] * inv * q[j * n + k];
  }
  return rank;
}

// Point-block inverse (cpu.cpp:218-262; sba_v1.cu:1016-1069): rank 3 -> inverse, rank 2 -> drop the
// smallest singular direction, rank <= 1 -> zero.
template <class T>
int point_pinv(const T M[9], T out[9]) {
  const int rank = sym_pinv<T>(3, M, T(1e-6), out);
  if (rank < 2) std::fill(out, out + 9, T(0));
  return rank;
}

// Lower Cholesky A = L L^T then L L^T x = b (the GPU bundler's spotrf/spotrs, libs/cuda_modules/sba.cpp:
// 606-640). Returns false when a pivot is not positive (devInfo != 0).
template <class T>
bool chol_solve(int n, const T* A, const T* b, T* x) {
  std::vector<T> L(A, A + n * n);
  for (int k = 0; k < n; ++k) {
    T d = L[k * n + k];
    for (int t = 0; t < k; ++t) d -= L[k * n + t] * L[k * n + t];
    if (!(d > T(0))) return false;
    const T lkk = std::sqrt(d);
    L[k * n + k] = lkk;
    for (int i = k + 1; i < n; ++i) {
      T s = L[i * n + k];
      for (int t = 0; t < k; ++t) s -= L[i * n + t] * L[k * n + t];
      L[i * n + k] = s / lkk;
    }
  }
  std::vector<T> y(n);
  for (int i = 0; i < n; ++i) {
    T s = b[i];
    for (int t = 0; t < i; ++t) s -= L[i * n + t] * y[t];
    y[i] = s / L[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    T s = y[i];
    for (int t = i + 1; t < n; ++t) s -= L[t * n + i] * x[t];
    x[i] = s / L[i * n + i];
  }
  return true;
}

// Closest rotation U V^T of a 3x3 matrix with the reflection fix, common/rotation_utils.h:36-50.
template <class T>
void nearest_rotation(const T M[3][3], T R[3][3]) {
  T MtM[9], ev[3], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) MtM[i * 3 + j] = M[0][i] * M[0][j] + M[1][i] * M[1][j] + M[2][i] * M[2][j];
  sym_jacobi<T>(3, MtM, ev, V);
  int order[3] = {0, 1, 2};  // singular values descending
  std::sort(order, order + 3, [&](int a, int b) { return ev[a] > ev[b]; });
  T U[3][3], Vs[3][3];
  for (int k = 0; k < 3; ++k) {
    const int c = order[k];
    const T s = std::sqrt(std::max(ev[c], T(0)));
    for (int i = 0; i < 3; ++i) Vs[i][k] = V[i * 3 + c];
    for (int i = 0; i < 3; ++i)
      U[i][k] = s > T(0) ? (M[i][0] * Vs[0][k] + M[i][1] * Vs[1][k] + M[i][2] * Vs[2][k]) / s : T(0);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = U[i][0] * Vs[j][0] + U[i][1] * Vs[j][1] + U[i][2] * Vs[j][2];
  const T det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) - R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
  if (det < T(0))
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[i][j] -= T(2) * U[i][2] * Vs[j][2];
}

template <class T>
struct Model {  // ModelFunction, schur_complement_bundler_cpu.h:29-41
  std::vector<T> r;   // 2 per observation
  std::vector<T> Jc;  // 2x6 row-major per observation
  std::vector<T> Jp;  // 2x3 row-major per observation
  std::vector<T> w;
};

template <class T>
struct Full {  // FullSystem, schur_complement_bundler_cpu.h:50-56 (block-sparse point_pose_block)
  int np, nc;
  std::vector<T> U;   // [nc][6][6]
  std::vector<T> ea;  // [6 nc]
  std::vector<T> V;   // [np][3][3]
  std::vector<T> eb;  // [3 np]
  // point_pose_block blocks hpc = w Jp^T L Jc (3x6) per (free point, free pose), listed per point
  std::vector<std::vector<std::pair<int, std::vector<T>>>> Wt;
};

template <class T>
struct Solver {
  // problem
  int np, npf, no, nc, ncf, ncam;
  std::vector<T> pts;
  std::vector<Pose<T>> poses;
  std::vector<Pose<T>> cams;
  std::vector<T> xy_, info_;  // observations in the solver's precision (info column-major 2x2)
  const OrcProblem* P;
  OrcOptions opt;
  T delta;

  // replaces the FP32 state of the problem by an FP64 one (points [np][3], poses [nc][16] column-major)
  void override_state(const double* p, const double* m) {
    if (p)
      for (size_t i = 0; i < pts.size(); ++i) pts[i] = static_cast<T>(p[i]);
    if (m)
      for (int j = 0; j < nc; ++j)
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) poses[j].R[r][c] = static_cast<T>(m[16 * j + c * 4 + r]);
          poses[j].t[r] = static_cast<T>(m[16 * j + 12 + r]);
        }
  }

  void load(const OrcProblem* p, const OrcOptions* o) {
    P = p;
    opt = *o;
    np = p->num_points;
    npf = p->num_points - p->num_fixed_points;
    no = p->num_observations;
    nc = p->num_poses;
    ncf = p->num_poses - p->num_fixed_key_frames;
    ncam = p->num_cameras;
    pts.resize(3 * static_cast<size_t>(np));
    for (size_t i = 0; i < pts.size(); ++i) pts[i] = static_cast<T>(p->points[i]);
    poses.resize(nc);
    for (int j = 0; j < nc; ++j) poses[j] = load_pose<T>(p->rig_from_world + 16 * j);
    cams.resize(ncam);
    for (int c = 0; c < ncam; ++c) cams[c] = load_pose<T>(p->camera_from_rig + 16 * c);
    delta = static_cast<T>(p->robustifier_scale);
    xy_.resize(2 * static_cast<size_t>(no));
    info_.resize(4 * static_cast<size_t>(no));
    for (size_t k = 0; k < xy_.size(); ++k) xy_[k] = static_cast<T>(p->observation_xys[k]);
    for (size_t k = 0; k < info_.size(); ++k) info_[k] = static_cast<T>(p->observation_infos[k]);
  }

  // replaces the observations by FP64 ones (xy [no][2], info [no][4] column-major)
  void override_observations(const double* xy, const double* inf) {
    if (xy)
      for (size_t k = 0; k < xy_.size(); ++k) xy_[k] = static_cast<T>(xy[k]);
    if (inf)
      for (size_t k = 0; k < info_.size(); ++k) info_[k] = static_cast<T>(inf[k]);
  }

  T info(int o, int r, int c) const { return info_[4 * o + c * 2 + r]; }

  // UpdateModel, cpu.cpp:64-131.
  void update_model(Model<T>& m) const {
    m.r.assign(2 * no, T(0));
    m.Jc.assign(12 * no, T(0));
    m.Jp.assign(6 * no, T(0));
    m.w.assign(no, T(0));
    for (int o = 0; o < no; ++o) {
      const int pi = P->point_ids[o], pj = P->pose_ids[o], cam = P->camera_ids[o];
      const Pose<T>& Tcr = cams[cam];
      const Pose<T> Tcw = compose(Tcr, poses[pj]);
      T pc[3];
      apply(Tcw, &pts[3 * pi], pc);
      if (!(pc[2] > T(kHither))) continue;
      const T pred[2] = {pc[0] / pc[2], pc[1] / pc[2]};
      const T r[2] = {xy_[2 * o] - pred[0],
                      xy_[2 * o + 1] - pred[1]};
      const T iz = T(1) / pc[2];
      const T dproj[2][3] = {{iz, T(0), -pred[0] * iz}, {T(0), iz, -pred[1] * iz}};
      T* Jp = &m.Jp[6 * o];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 3; ++b)
          Jp[a * 3 + b] = dproj[a][0] * Tcw.R[0][b] + dproj[a][1] * Tcw.R[1][b] + dproj[a][2] * Tcw.R[2][b];
      // camera twist Jacobian [dproj * skew(-pc), dproj], then rig twist via Adjoint(camera_from_rig)
      const T S[3][3] = {{T(0), pc[2], -pc[1]}, {-pc[2], T(0), pc[0]}, {pc[1], -pc[0], T(0)}};  // skew(-pc)
      T Jcam[2][6];
      for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 3; ++b) Jcam[a][b] = dproj[a][0] * S[0][b] + dproj[a][1] * S[1][b] + dproj[a][2] * S[2][b];
        for (int b = 0; b < 3; ++b) Jcam[a][3 + b] = dproj[a][b];
      }
      T Ad[6][6] = {};
      const T tx[3][3] = {{T(0), -Tcr.t[2], Tcr.t[1]}, {Tcr.t[2], T(0), -Tcr.t[0]}, {-Tcr.t[1], Tcr.t[0], T(0)}};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          Ad[a][b] = Tcr.R[a][b];
          Ad[3 + a][3 + b] = Tcr.R[a][b];
          Ad[3 + a][b] = tx[a][0] * Tcr.R[0][b] + tx[a][1] * Tcr.R[1][b] + tx[a][2] * Tcr.R[2][b];
        }
      T* Jc = &m.Jc[12 * o];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 6; ++b) {
          T s = 0;
          for (int k = 0; k < 6; ++k) s += Jcam[a][k] * Ad[k][b];
          Jc[a * 6 + b] = s;
        }
      m.r[2 * o] = r[0];
      m.r[2 * o + 1] = r[1];
      const T sq = r[0] * (info(o, 0, 0) * r[0] + info(o, 0, 1) * r[1]) + r[1] * (info(o, 1, 0) * r[0] + info(o, 1, 1) * r[1]);
      m.w[o] = loss_drho<T>(opt.loss, sq, delta);
    }
  }

  // BuildFullSystem, cpu.cpp:136-209.
  void build_full(Full<T>& F, const Model<T>& m) const {
    F.np = npf;
    F.nc = ncf;
    F.U.assign(36 * static_cast<size_t>(ncf), T(0));
    F.ea.assign(6 * static_cast<size_t>(ncf), T(0));
    F.V.assign(9 * static_cast<size_t>(npf), T(0));
    F.eb.assign(3 * static_cast<size_t>(npf), T(0));
    F.Wt.assign(npf, {});
    for (int o = 0; o < no; ++o) {
      const T w = m.w[o];
      const T* Jc = &m.Jc[12 * o];
      const T* Jp = &m.Jp[6 * o];
      T L[2][2] = {{info(o, 0, 0), info(o, 0, 1)}, {info(o, 1, 0), info(o, 1, 1)}};
      T LJc[2][6], LJp[2][3];
      for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 6; ++b) LJc[a][b] = L[a][0] * Jc[b] + L[a][1] * Jc[6 + b];
        for (int b = 0; b < 3; ++b) LJp[a][b] = L[a][0] * Jp[b] + L[a][1] * Jp[3 + b];
      }
      const T wr[2] = {w * m.r[2 * o], w * m.r[2 * o + 1]};
      const T Lwr[2] = {L[0][0] * wr[0] + L[0][1] * wr[1], L[1][0] * wr[0] + L[1][1] * wr[1]};
      const int c = P->pose_ids[o], p = P->point_ids[o];
      if (c < ncf && p < npf) {
        auto& lst = F.Wt[p];
        std::vector<T>* blk = nullptr;
        for (auto& e : lst)
          if (e.first == c) blk = &e.second;
        if (!blk) {
          lst.emplace_back(c, std::vector<T>(18, T(0)));
          blk = &lst.back().second;
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 6; ++b) (*blk)[a * 6 + b] += (Jp[a] * LJc[0][b] + Jp[3 + a] * LJc[1][b]) * w;
      }
      if (c < ncf) {
        T* U = &F.U[36 * c];
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b) U[a * 6 + b] += (Jc[a] * LJc[0][b] + Jc[6 + a] * LJc[1][b]) * w;
        for (int a = 0; a < 6; ++a) F.ea[6 * c + a] += Jc[a] * Lwr[0] + Jc[6 + a] * Lwr[1];
      }
      if (p < npf) {
        T* V = &F.V[9 * p];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) V[a * 3 + b] += (Jp[a] * LJp[0][b] + Jp[3 + a] * LJp[1][b]) * w;
        for (int a = 0; a < 3; ++a) F.eb[3 * p + a] += Jp[a] * Lwr[0] + Jp[3 + a] * Lwr[1];
      }
    }
  }

  struct Reduced {  // ReducedSystem, cpu.h:58-64 (Y kept block-sparse per point)
    std::vector<T> S, e, inv, gb;
    std::vector<std::vector<std::pair<int, std::vector<T>>>> Y;  // per point: (pose, 6x3 block)
  };

  // BuildReducedSystem, cpu.cpp:211-287.
  void build_reduced(Reduced& R, const Full<T>& F, T lambda) const {
    const int n = 6 * ncf;
    R.inv.assign(9 * static_cast<size_t>(npf), T(0));
    R.gb.assign(3 * static_cast<size_t>(npf), T(0));
    for (int i = 0; i < npf; ++i) {
      T M[9];
      for (int k = 0; k < 9; ++k) M[k] = F.V[9 * i + k];
      for (int d = 0; d < 3; ++d) M[4 * d] += F.V[9 * i + 4 * d] * lambda;
      point_pinv<T>(M, &R.inv[9 * i]);
      for (int a = 0; a < 3; ++a)
        R.gb[3 * i + a] = R.inv[9 * i + 3 * a] * F.eb[3 * i] + R.inv[9 * i + 3 * a + 1] * F.eb[3 * i + 1] +
                          R.inv[9 * i + 3 * a + 2] * F.eb[3 * i + 2];
    }
    R.S.assign(static_cast<size_t>(n) * n, T(0));
    for (int c = 0; c < ncf; ++c)
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) R.S[(6 * c + a) * n + 6 * c + b] = F.U[36 * c + 6 * a + b];
    R.e.assign(F.ea.begin(), F.ea.end());
    R.Y.assign(npf, {});
    if (opt.dense_schur) {
      // temp = point_pose_block^T (n x 3np), column blocks *= inv, S -= temp * point_pose_block (cpu.cpp:264-277)
      const size_t K = 3 * static_cast<size_t>(npf);
      std::vector<T> ppb(K * n, T(0));  // point_pose_block, row-major (3np x n)
      for (int i = 0; i < npf; ++i)
        for (auto& e : F.Wt[i])
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 6; ++b) ppb[(3 * i + a) * n + 6 * e.first + b] = e.second[a * 6 + b];
      std::vector<T> temp(static_cast<size_t>(n) * K);  // row-major (n x 3np)
      for (size_t k = 0; k < K; ++k)
        for (int a = 0; a < n; ++a) temp[a * K + k] = ppb[k * n + a];
      for (int i = 0; i < npf; ++i)
        for (int a = 0; a < n; ++a) {
          T row[3] = {temp[a * K + 3 * i], temp[a * K + 3 * i + 1], temp[a * K + 3 * i + 2]};
          for (int c = 0; c < 3; ++c)
            temp[a * K + 3 * i + c] = row[0] * R.inv[9 * i + c] + row[1] * R.inv[9 * i + 3 + c] + row[2] * R.inv[9 * i + 6 + c];
        }
      for (int a = 0; a < n; ++a) {
        T* Srow = &R.S[static_cast<size_t>(a) * n];
        const T* trow = &temp[a * K];
        for (size_t k = 0; k < K; ++k) {
          const T t = trow[k];
          const T* prow = &ppb[k * n];
          for (int b = 0; b < n; ++b) Srow[b] -= t * prow[b];
        }
        T s = 0;
        for (size_t k = 0; k < K; ++k) s += trow[k] * F.eb[k];
        R.e[a] -= s;
      }
      for (int i = 0; i < npf; ++i)
        for (auto& e : F.Wt[i]) {
          std::vector<T> y(18);
          for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 3; ++c) y[a * 3 + c] = temp[(6 * e.first + a) * K + 3 * i + c];
          R.Y[i].emplace_back(e.first, y);
        }
    } else {
      for (int i = 0; i < npf; ++i) {
        const T* inv = &R.inv[9 * i];
        for (auto& e : F.Wt[i]) {
          std::vector<T> y(18);  // Y = hpc^T * inv (6x3)
          for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 3; ++c)
              y[a * 3 + c] = e.second[a] * inv[c] + e.second[6 + a] * inv[3 + c] + e.second[12 + a] * inv[6 + c];
          R.Y[i].emplace_back(e.first, y);
        }
        for (auto& ya : R.Y[i]) {
          const int ca = ya.first;
          for (auto& wb : F.Wt[i]) {
            const int cb = wb.first;
            for (int a = 0; a < 6; ++a)
              for (int b = 0; b < 6; ++b)
                R.S[(6 * ca + a) * n + 6 * cb + b] -= ya.second[a * 3] * wb.second[b] + ya.second[a * 3 + 1] * wb.second[6 + b] +
                                                      ya.second[a * 3 + 2] * wb.second[12 + b];
          }
          for (int a = 0; a < 6; ++a)
            R.e[6 * ca + a] -= ya.second[a * 3] * F.eb[3 * i] + ya.second[a * 3 + 1] * F.eb[3 * i + 1] +
                               ya.second[a * 3 + 2] * F.eb[3 * i + 2];
        }
      }
    }
    if (opt.round_reduced_f32 == 2) {
      // S currently holds U - sum W V*^-1 W^T and e holds e_a - sum W V*^-1 e_b: split off the Schur sums, round
      // both parts to FP32, and combine again in FP64
      auto r32 = [](T x) { return static_cast<T>(static_cast<float>(x)); };
      std::vector<T> U(static_cast<size_t>(n) * n, T(0));
      for (int c = 0; c < ncf; ++c)
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b) U[(6 * c + a) * n + 6 * c + b] = F.U[36 * c + 6 * a + b];
      for (size_t k = 0; k < U.size(); ++k) R.S[k] = r32(U[k]) - r32(U[k] - R.S[k]);
      for (int k = 0; k < n; ++k) R.e[k] = r32(F.ea[k]) - r32(F.ea[k] - R.e[k]);
    }
    for (int i = 0; i < n; ++i) R.S[static_cast<size_t>(i) * n + i] += lambda * F.U[36 * (i / 6) + 7 * (i % 6)];
    if (opt.round_reduced_f32 == 1) {
      for (auto& x : R.S) x = static_cast<T>(static_cast<float>(x));
      for (auto& x : R.e) x = static_cast<T>(static_cast<float>(x));
    }
  }

  struct Update {
    std::vector<T> dc, dp;  // pose_step (6 ncf), point_step (3 npf)
    std::vector<Pose<T>> pose;
    bool ok = true;
  };

  // ComputeUpdate, cpu.cpp:337-365 (Jacobi pseudo-inverse) or the GPU bundler's Cholesky (sba.cpp:606-640).
  void compute_update(Update& u, const Reduced& R) const {
    const int n = 6 * ncf;
    u.dc.assign(n, T(0));
    u.ok = true;
    if (n > 0) {
      if (opt.rcs_solver == 0) {
        u.ok = chol_solve<T>(n, R.S.data(), R.e.data(), u.dc.data());
      } else {
        std::vector<T> pinv(static_cast<size_t>(n) * n);
        sym_pinv<T>(n, R.S.data(), T(1e-6), pinv.data());
        for (int a = 0; a < n; ++a) {
          T s = 0;
          for (int b = 0; b < n; ++b) s += pinv[static_cast<size_t>(a) * n + b] * R.e[b];
          u.dc[a] = s;
        }
      }
    }
    u.dp.assign(3 * static_cast<size_t>(npf), T(0));
    for (int i = 0; i < npf; ++i) {
      T d[3] = {R.gb[3 * i], R.gb[3 * i + 1], R.gb[3 * i + 2]};
      for (auto& y : R.Y[i])
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 6; ++a) d[c] -= y.second[a * 3 + c] * u.dc[6 * y.first + a];
      for (int c = 0; c < 3; ++c) u.dp[3 * i + c] = d[c];
    }
    u.pose.resize(ncf);
    for (int j = 0; j < ncf; ++j) u.pose[j] = twist_exp<T>(&u.dc[6 * j]);
  }

  // EvaluateCost, cpu.cpp:289-335. Returns +inf when every observation is inside the near plane.
  T evaluate_cost(const Update* u, int* skipped_out) const {
    T cost = 0;
    int skipped = 0;
    for (int o = 0; o < no; ++o) {
      const int pi = P->point_ids[o], pj = P->pose_ids[o], cam = P->camera_ids[o];
      T pw[3] = {pts[3 * pi], pts[3 * pi + 1], pts[3 * pi + 2]};
      if (u && pi < npf)
        for (int d = 0; d < 3; ++d) pw[d] += u->dp[3 * pi + d];
      T pr0[3], pr[3], pc[3];
      apply(poses[pj], pw, pr0);
      if (u && pj < ncf)
        apply(u->pose[pj], pr0, pr);
      else
        std::copy(pr0, pr0 + 3, pr);
      apply(cams[cam], pr, pc);
      if (pc[2] > T(kHither)) {
        const T r[2] = {xy_[2 * o] - pc[0] / pc[2],
                        xy_[2 * o + 1] - pc[1] / pc[2]};
        const T sq = r[0] * (info(o, 0, 0) * r[0] + info(o, 0, 1) * r[1]) + r[1] * (info(o, 1, 0) * r[0] + info(o, 1, 1) * r[1]);
        cost += loss_rho<T>(opt.loss, sq, delta);
      } else {
        ++skipped;
      }
    }
    if (skipped_out) *skipped_out = skipped;
    if (skipped == no) return std::numeric_limits<T>::infinity();
    return cost;
  }

  // UpdateState, cpu.cpp:367-383.
  void update_state(const Update& u) {
    for (int i = 0; i < npf; ++i)
      for (int d = 0; d < 3; ++d) pts[3 * i + d] += u.dp[3 * i + d];
    for (int j = 0; j < ncf; ++j) {
      Pose<T> m = compose(u.pose[j], poses[j]);
      if (opt.reorthonormalize) {
        T R[3][3];
        nearest_rotation<T>(m.R, R);
        std::memcpy(m.R, R, sizeof(R));
      }
      poses[j] = m;
    }
  }

  // ComputePredictedRelativeReduction, cpu.cpp:385-433; returns the numerators (hessian, scaling) too.
  T predicted(T current, T lambda, const Update& u, const Full<T>& F, T* hess_out, T* scal_out) const {
    T pose_h = 0, cross = 0, point_h = 0, pose_s = 0, point_s = 0;
    for (int c = 0; c < ncf; ++c)
      for (int a = 0; a < 6; ++a) {
        T s = 0;
        for (int b = 0; b < 6; ++b) s += F.U[36 * c + 6 * a + b] * u.dc[6 * c + b];
        pose_h += u.dc[6 * c + a] * s;
        pose_s += u.dc[6 * c + a] * F.U[36 * c + 7 * a] * u.dc[6 * c + a];
      }
    for (int i = 0; i < npf; ++i) {
      for (auto& e : F.Wt[i])
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 6; ++b) cross += u.dp[3 * i + a] * e.second[a * 6 + b] * u.dc[6 * e.first + b];
      for (int a = 0; a < 3; ++a) {
        T s = 0;
        for (int b = 0; b < 3; ++b) s += F.V[9 * i + 3 * a + b] * u.dp[3 * i + b];
        point_h += u.dp[3 * i + a] * s;
        point_s += u.dp[3 * i + a] * F.V[9 * i + 4 * a] * u.dp[3 * i + a];
      }
    }
    const T hess = pose_h + T(2) * cross + point_h;
    const T scal = pose_s + point_s;
    if (hess_out) *hess_out = hess;
    if (scal_out) *scal_out = scal;
    return (pose_h + cross) / current + (cross + point_h) / current + ((pose_s + point_s) / current) * T(2) * lambda;
  }

  static T inf_norm(const std::vector<T>& v) {
    T m = 0;
    for (T x : v) m = std::max(m, std::fabs(x));
    return m;
  }

  // SchurComplementBundlerCpu::solve (cpu.cpp:435-529) with the GPU bundler's handling of a failed
  // factorization (gpu.cpp:204-209).
  int solve(OrcResult* res) {
    res->iterations = 0;
    res->accepted = 0;
    if (np == 0 || nc == 0 || no == 0) return 1;  // gpu.cpp:121-123
    Update u;
    u.dc.assign(6 * static_cast<size_t>(ncf), T(0));
    u.dp.assign(3 * static_cast<size_t>(npf), T(0));
    Model<T> model;
    update_model(model);
    const T initial = evaluate_cost(nullptr, nullptr);
    T current = initial;
    res->initial_cost = static_cast<double>(initial);
    res->last_cost = static_cast<double>(initial);
    res->lambda = 1e-3;
    if (initial < T(kFltEps)) return 0;
    if (std::isinf(initial)) return 2;
    Full<T> F;
    build_full(F, model);
    T lambda = T(0.001f);
    Reduced R;
    int it = 0;
    while (it < P->max_iterations) {
      ++it;
      res->iterations = it;
      int flags = 0;
      build_reduced(R, F, lambda);
      compute_update(u, R);
      const T lambda_used = lambda;
      T cost = 0, rho = 0;
      bool stop = false;
      if (!u.ok) {
        lambda *= T(5);
        flags |= 2;
      } else {
        cost = evaluate_cost(&u, nullptr);
        if (current < initial * T(kFltEps)) {
          stop = true;
          flags |= 4;
        } else {
          const T pred = predicted(current, lambda, u, F, nullptr, nullptr);
          if (!opt.fixed_iterations && pred < T(kSqrtFltEps) && inf_norm(u.dc) < T(kSqrtFltEps) &&
              inf_norm(u.dp) < T(kSqrtFltEps)) {
            stop = true;
            flags |= 4;
          } else {
            rho = (T(1) - cost / current) / pred;
            bool accept = rho > T(0.25f), shrink = rho > T(0.75f);
            if (it == opt.flip_iteration) {
              if (std::fabs(rho - T(0.75f)) < std::fabs(rho - T(0.25f)))
                shrink = !shrink;
              else
                accept = !accept;
            }
            if (accept) {
              if (shrink && lambda * T(0.125f) > T(0)) lambda *= T(0.125f);
              current = cost;
              res->last_cost = static_cast<double>(cost);
              res->accepted += 1;
              flags |= 1;
              update_state(u);
              update_model(model);
              build_full(F, model);
            } else {
              lambda *= T(5);
            }
          }
        }
      }
      if (res->trace_capacity >= it) {
        if (res->trace_lambda) res->trace_lambda[it - 1] = static_cast<double>(lambda_used);
        if (res->trace_cost) res->trace_cost[it - 1] = static_cast<double>(cost);
        if (res->trace_rho) res->trace_rho[it - 1] = static_cast<double>(rho);
        if (res->trace_flags) res->trace_flags[it - 1] = flags;
      }
      if (stop) break;
    }
    res->lambda = static_cast<double>(lambda);
    return 0;
  }

  void export_state(OrcResult* res) const {
    if (res->points)
      for (int i = 0; i < 3 * np; ++i) res->points[i] = static_cast<double>(pts[i]);
    if (res->poses)
      for (int j = 0; j < nc; ++j) store_pose(poses[j], res->poses + 16 * j);
  }
};

// ---------------------------------------------------------------------------------------------------------
// Implicit gradient of a loss of the converged state (no reference counterpart; the reference's analytic
// Jacobians of cpu.cpp:101-118 and the information conversion of camera/observation.cpp:57-67 are the anchors).
// At a solution x* the gradient of F(x) = sum rho(r^T L r) vanishes, so for parameters theta of the
// observations (xy, L): dx*/dtheta = -H^-1 d(grad F)/dtheta, with H the full Hessian of F at x*, and
// dL/dtheta = -a^T d(grad F)/dtheta where H a = dL/dx. H includes the rho'' terms and the second-order terms
// of the projection and of the left-multiplied twist; H a = g is solved by preconditioned conjugate gradients,
// the preconditioner being the Gauss-Newton system 2 sum rho' J^T L J solved through its Schur complement
// (lambda = 0). Tangent coordinates: left twists (omega, v) of the free key frames, then the free points.
struct ObsLin {
  double r[2], J[2][9], L[2][2], w, w2, q0[3], Rcr[3][3], Rrw[3][3], P[3];
  int pose, point;
  bool valid;
};

double loss_d2rho(int kind, double s, double d) {
  if (kind == 0) return -(d + 1.0) / ((d + s) * (d + s));
  if (d <= 0.0 || s < d * d) return 0.0;
  return -0.25 * d / (s * std::sqrt(s));
}

struct Adjoint {
  Solver<double>& S;
  int nx, ncf, npf;
  std::vector<ObsLin> obs;
  Full<double> F;
  Solver<double>::Reduced R;

  explicit Adjoint(Solver<double>& s) : S(s) {
    ncf = S.ncf;
    npf = S.npf;
    nx = 6 * ncf + 3 * npf;
    obs.resize(S.no);
    for (int o = 0; o < S.no; ++o) {
      ObsLin& L = obs[o];
      L.pose = S.P->pose_ids[o];
      L.point = S.P->point_ids[o];
      const int cam = S.P->camera_ids[o];
      const Pose<double>& Tcr = S.cams[cam];
      const Pose<double>& Trw = S.poses[L.pose];
      apply(Trw, &S.pts[3 * L.point], L.q0);
      apply(Tcr, L.q0, L.P);
      std::memcpy(L.Rcr, Tcr.R, sizeof(L.Rcr));
      std::memcpy(L.Rrw, Trw.R, sizeof(L.Rrw));
      L.valid = L.P[2] > double(kHither);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) L.L[a][b] = S.info(o, a, b);
      std::memset(L.J, 0, sizeof(L.J));
      L.w = L.w2 = 0;
      L.r[0] = L.r[1] = 0;
      if (!L.valid) continue;
      const double iz = 1.0 / L.P[2], u = L.P[0] * iz, v = L.P[1] * iz;
      L.r[0] = S.xy_[2 * o] - u;
      L.r[1] = S.xy_[2 * o + 1] - v;
      const double dproj[2][3] = {{iz, 0, -u * iz}, {0, iz, -v * iz}};
      // D1 = R_cr [-[q0]x, I, R_rw]
      const double mq[3][3] = {{0, L.q0[2], -L.q0[1]}, {-L.q0[2], 0, L.q0[0]}, {L.q0[1], -L.q0[0], 0}};
      double D1[3][9];
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) {
          double s1 = 0, s3 = 0;
          for (int m = 0; m < 3; ++m) {
            s1 += L.Rcr[a][m] * mq[m][k];
            s3 += L.Rcr[a][m] * L.Rrw[m][k];
          }
          D1[a][k] = s1;
          D1[a][3 + k] = L.Rcr[a][k];
          D1[a][6 + k] = s3;
        }
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 9; ++k) L.J[a][k] = dproj[a][0] * D1[0][k] + dproj[a][1] * D1[1][k] + dproj[a][2] * D1[2][k];
      const double Lr0 = L.L[0][0] * L.r[0] + L.L[0][1] * L.r[1], Lr1 = L.L[1][0] * L.r[0] + L.L[1][1] * L.r[1];
      const double sq = L.r[0] * Lr0 + L.r[1] * Lr1;
      L.w = loss_drho<double>(S.opt.loss, sq, S.delta);
      L.w2 = loss_d2rho(S.opt.loss, sq, S.delta);
    }
    Model<double> m;
    S.update_model(m);
    S.build_full(F, m);
    S.build_reduced(R, F, 0.0);
  }

  // gather the observation's 9 tangent entries of a global vector (zero for fixed entries)
  void gather(const ObsLin& L, const double* x, double v[9]) const {
    for (int k = 0; k < 9; ++k) v[k] = 0;
    if (L.pose < ncf)
      for (int k = 0; k < 6; ++k) v[k] = x[6 * L.pose + k];
    if (L.point < npf)
      for (int k = 0; k < 3; ++k) v[6 + k] = x[6 * ncf + 3 * L.point + k];
  }
  void scatter(const ObsLin& L, const double h[9], double* y) const {
    if (L.pose < ncf)
      for (int k = 0; k < 6; ++k) y[6 * L.pose + k] += h[k];
    if (L.point < npf)
      for (int k = 0; k < 3; ++k) y[6 * ncf + 3 * L.point + k] += h[6 + k];
  }

  // y = H x with the full Hessian of F
  void hvp(const double* x, double* y) const {
    std::fill(y, y + nx, 0.0);
    for (const ObsLin& L : obs) {
      if (!L.valid) continue;
      double v[9];
      gather(L, x, v);
      double u[2] = {0, 0};
      for (int k = 0; k < 9; ++k) {
        u[0] += L.J[0][k] * v[k];
        u[1] += L.J[1][k] * v[k];
      }
      const double Lu[2] = {L.L[0][0] * u[0] + L.L[0][1] * u[1], L.L[1][0] * u[0] + L.L[1][1] * u[1]};
      const double Lr[2] = {L.L[0][0] * L.r[0] + L.L[0][1] * L.r[1], L.L[1][0] * L.r[0] + L.L[1][1] * L.r[1]};
      const double rLu = L.r[0] * Lu[0] + L.r[1] * Lu[1];
      double h[9];
      for (int k = 0; k < 9; ++k) {
        const double JLr = L.J[0][k] * Lr[0] + L.J[1][k] * Lr[1];
        h[k] = 2 * L.w * (L.J[0][k] * Lu[0] + L.J[1][k] * Lu[1]) + 4 * L.w2 * JLr * rLu;
      }
      // - 2 Hess(psi) v with psi = c^T pi(p_c), c = rho' L r
      const double c[2] = {L.w * Lr[0], L.w * Lr[1]};
      const double iz = 1.0 / L.P[2];
      const double g3[3] = {c[0] * iz, c[1] * iz, -(c[0] * L.P[0] + c[1] * L.P[1]) * iz * iz};
      const double Hp[3][3] = {{0, 0, -c[0] * iz * iz},
                               {0, 0, -c[1] * iz * iz},
                               {-c[0] * iz * iz, -c[1] * iz * iz, 2 * (c[0] * L.P[0] + c[1] * L.P[1]) * iz * iz * iz}};
      // D1 v and D1^T Hp (D1 v)
      const double mq[3][3] = {{0, L.q0[2], -L.q0[1]}, {-L.q0[2], 0, L.q0[0]}, {L.q0[1], -L.q0[0], 0}};
      double D1[3][9];
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) {
          double s1 = 0, s3 = 0;
          for (int m = 0; m < 3; ++m) {
            s1 += L.Rcr[a][m] * mq[m][k];
            s3 += L.Rcr[a][m] * L.Rrw[m][k];
          }
          D1[a][k] = s1;
          D1[a][3 + k] = L.Rcr[a][k];
          D1[a][6 + k] = s3;
        }
      double dv[3] = {0, 0, 0};
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 9; ++k) dv[a] += D1[a][k] * v[k];
      double hd[3];
      for (int a = 0; a < 3; ++a) hd[a] = Hp[a][0] * dv[0] + Hp[a][1] * dv[1] + Hp[a][2] * dv[2];
      double hpsi[9];
      for (int k = 0; k < 9; ++k) hpsi[k] = D1[0][k] * hd[0] + D1[1][k] * hd[1] + D1[2][k] * hd[2];
      // second-order part of p_r along v: g' = R_cr^T g3
      double gp[3];
      for (int m = 0; m < 3; ++m) gp[m] = L.Rcr[0][m] * g3[0] + L.Rcr[1][m] * g3[1] + L.Rcr[2][m] * g3[2];
      const double* wv = v;
      const double* vv = v + 3;
      double Rdp[3];
      for (int m = 0; m < 3; ++m) Rdp[m] = L.Rrw[m][0] * v[6] + L.Rrw[m][1] * v[7] + L.Rrw[m][2] * v[8];
      auto cross = [](const double a[3], const double b[3], double o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
      };
      double wxq[3], avec[3], gxw[3], t1[3], t2[3];
      cross(wv, L.q0, wxq);
      for (int m = 0; m < 3; ++m) avec[m] = 0.5 * wxq[m] + 0.5 * vv[m] + Rdp[m];
      cross(avec, gp, t1);      // grad_omega of g'.(omega x a)
      cross(gp, wv, gxw);       // g' x omega_v
      cross(L.q0, gxw, t2);     // grad_omega of g'.(omega_v x (omega x q0)) / 2 * 2
      for (int m = 0; m < 3; ++m) hpsi[m] += t1[m] + 0.5 * t2[m];
      for (int m = 0; m < 3; ++m) hpsi[3 + m] += 0.5 * gxw[m];
      for (int m = 0; m < 3; ++m) hpsi[6 + m] += L.Rrw[0][m] * gxw[0] + L.Rrw[1][m] * gxw[1] + L.Rrw[2][m] * gxw[2];
      if (L.pose >= ncf)
        for (int k = 0; k < 6; ++k) hpsi[k] = 0;
      if (L.point >= npf)
        for (int k = 0; k < 3; ++k) hpsi[6 + k] = 0;
      for (int k = 0; k < 9; ++k) h[k] -= 2 * hpsi[k];
      scatter(L, h, y);
    }
  }

  // z = M^-1 q with M = 2 sum rho' J^T L J (the LM system at lambda = 0, doubled), via its Schur complement
  bool precond(const double* q, double* z) const {
    const int n = 6 * ncf;
    std::vector<double> e(q, q + n), zc(n, 0.0);
    std::vector<double> t(3 * static_cast<size_t>(npf));
    for (int i = 0; i < npf; ++i)
      for (int a = 0; a < 3; ++a)
        t[3 * i + a] = R.inv[9 * i + 3 * a] * q[n + 3 * i] + R.inv[9 * i + 3 * a + 1] * q[n + 3 * i + 1] +
                       R.inv[9 * i + 3 * a + 2] * q[n + 3 * i + 2];
    for (int i = 0; i < npf; ++i)
      for (auto& y : R.Y[i])
        for (int a = 0; a < 6; ++a)
          e[6 * y.first + a] -= y.second[3 * a] * q[n + 3 * i] + y.second[3 * a + 1] * q[n + 3 * i + 1] +
                                y.second[3 * a + 2] * q[n + 3 * i + 2];
    if (n > 0 && !chol_solve<double>(n, R.S.data(), e.data(), zc.data())) return false;
    for (int k = 0; k < n; ++k) z[k] = 0.5 * zc[k];
    for (int i = 0; i < npf; ++i) {
      double d[3] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
      for (auto& y : R.Y[i])
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 6; ++a) d[c] -= y.second[a * 3 + c] * zc[6 * y.first + a];
      for (int c = 0; c < 3; ++c) z[n + 3 * i + c] = 0.5 * d[c];
    }
    return true;
  }

  // preconditioned conjugate gradients on H a = g; returns iterations used (negative on failure)
  int pcg(const double* g, double* a, int max_it, double tol) const {
    std::vector<double> r(g, g + nx), z(nx), p(nx), q(nx);
    std::fill(a, a + nx, 0.0);
    if (!precond(r.data(), z.data())) return -1;
    p = z;
    double rz = 0, gg = 0;
    for (int k = 0; k < nx; ++k) {
      rz += r[k] * z[k];
      gg += g[k] * g[k];
    }
    if (gg == 0) return 0;
    for (int it = 1; it <= max_it; ++it) {
      hvp(p.data(), q.data());
      double pq = 0;
      for (int k = 0; k < nx; ++k) pq += p[k] * q[k];
      if (!(pq > 0)) return it - 1;  // non-positive curvature: stop with the current a (truncated CG)
      const double alpha = rz / pq;
      double rr = 0;
      for (int k = 0; k < nx; ++k) {
        a[k] += alpha * p[k];
        r[k] -= alpha * q[k];
        rr += r[k] * r[k];
      }
      if (std::sqrt(rr / gg) < tol) return it;
      precond(r.data(), z.data());
      double rz2 = 0;
      for (int k = 0; k < nx; ++k) rz2 += r[k] * z[k];
      const double beta = rz2 / rz;
      rz = rz2;
      for (int k = 0; k < nx; ++k) p[k] = z[k] + beta * p[k];
    }
    return max_it;
  }

  // dL/dxy (2 per observation) and dL/dL (3 per observation: L00, symmetric L01, L11) from the adjoint a
  void observation_grads(const double* a, double* dxy, double* dinfo) const {
    for (int o = 0; o < S.no; ++o) {
      const ObsLin& L = obs[o];
      double* gx = dxy + 2 * o;
      double* gi = dinfo + 3 * o;
      gx[0] = gx[1] = gi[0] = gi[1] = gi[2] = 0;
      if (!L.valid) continue;
      double v[9];
      gather(L, a, v);
      double Ja[2] = {0, 0};
      for (int k = 0; k < 9; ++k) {
        Ja[0] += L.J[0][k] * v[k];
        Ja[1] += L.J[1][k] * v[k];
      }
      const double LJa[2] = {L.L[0][0] * Ja[0] + L.L[0][1] * Ja[1], L.L[1][0] * Ja[0] + L.L[1][1] * Ja[1]};
      const double Lr[2] = {L.L[0][0] * L.r[0] + L.L[0][1] * L.r[1], L.L[1][0] * L.r[0] + L.L[1][1] * L.r[1]};
      const double rLJa = L.r[0] * LJa[0] + L.r[1] * LJa[1];
      gx[0] = 2 * (L.w * LJa[0] + 2 * L.w2 * rLJa * Lr[0]);
      gx[1] = 2 * (L.w * LJa[1] + 2 * L.w2 * rLJa * Lr[1]);
      // G = 2 [rho' (J a) r^T + rho'' ((J a)^T L r) r r^T]
      double G[2][2];
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) G[p][q] = 2 * (L.w * Ja[p] * L.r[q] + L.w2 * rLJa * L.r[p] * L.r[q]);
      gi[0] = G[0][0];
      gi[1] = G[0][1] + G[1][0];
      gi[2] = G[1][1];
    }
  }
};

template <class T>
int solve_impl(const OrcProblem* p, const OrcOptions* o, OrcResult* r, const double* sp = nullptr,
               const double* sq = nullptr, const double* oxy = nullptr, const double* oinfo = nullptr) {
  if (!p || !o || !r) return 5;
  Solver<T> s;
  s.load(p, o);
  s.override_state(sp, sq);
  s.override_observations(oxy, oinfo);
  const int status = s.solve(r);
  r->status = status;
  s.export_state(r);
  return status;
}


// ---------------------------------------------------------------------------------------------------------
// MonoSBASolver (libs/sba/mono_sba_solver.{h,cpp}): Hartley & Zisserman 2nd ed. Algorithm A6.4 with cameras
// parameterized by the twist a of camera_from_world, the Huber loss with per-track information matrices
// estimated from the track's residuals, multiplicative damping (1 + lambda) of the U and V diagonals, the x10 / /10
// damping schedule, the mean robust residual as the step test, and a pose covariance from the last free frame's
// block of the reduced system. Restated step for step, templated on the scalar (float: the reference's arithmetic).
// ---------------------------------------------------------------------------------------------------------

constexpr float kMonoLambdaMax = 1.0e+7f;  // mono_sba_solver.cpp:26
constexpr float kMonoHuber = 0.5f;         // mono_sba_solver.cpp:214

// Log of a pose (math/twist.cpp:44-67): the rotation axis is the right singular vector of R - I of its smallest
// singular value (an eigenvector of (R - I)^T (R - I)), the angle atan2(rvec . axis, trace(R) - 1).
template <class T>
void twist_log(const Pose<T>& P, T tw[6]) {
  T a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = P.R[i][j] - (i == j ? T(1) : T(0));
  T ata[9], ev[3], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ata[i * 3 + j] = a[0][i] * a[0][j] + a[1][i] * a[1][j] + a[2][i] * a[2][j];
  sym_jacobi<T>(3, ata, ev, V);
  int k = 0;
  for (int c = 1; c < 3; ++c)
    if (ev[c] < ev[k]) k = c;
  T w[3] = {V[0 * 3 + k], V[1 * 3 + k], V[2 * 3 + k]};
  const T nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (auto& x : w) x /= nw;
  const T rvec[3] = {a[2][1] - a[1][2], a[0][2] - a[2][0], a[1][0] - a[0][1]};
  const T wmag = std::atan2(rvec[0] * w[0] + rvec[1] * w[1] + rvec[2] * w[2], a[0][0] + a[1][1] + a[2][2] + T(2));
  const T h = wmag / T(2);
  const T* t = P.t;
  const T wxt[3] = {w[1] * t[2] - w[2] * t[1], w[2] * t[0] - w[0] * t[2], w[0] * t[1] - w[1] * t[0]};
  const T wxwxt[3] = {w[1] * wxt[2] - w[2] * wxt[1], w[2] * wxt[0] - w[0] * wxt[2], w[0] * wxt[1] - w[1] * wxt[0]};
  const T coef = std::fabs(wmag) < T(kFltEps) ? T(0) : (T(1) - h / std::tan(h));
  for (int i = 0; i < 3; ++i) {
    tw[i] = wmag * w[i];
    tw[3 + i] = t[i] + coef * wxwxt[i] - h * wxt[i];
  }
}

// VectorTwist::operator+ (math/twist.h:67-73): Log(Exp(rhs) Exp(lhs)).
template <class T>
void twist_add(const T lhs[6], const T rhs[6], T out[6]) {
  twist_log<T>(compose(twist_exp<T>(rhs), twist_exp<T>(lhs)), out);
}

template <class T>
T avoid_zero(T a) {  // common/types.h:219-221
  return std::fabs(a) > T(kFltEps) ? a : std::copysign(T(kFltEps), a);
}

// Eigen's computeInverseAndDetWithCheck for 2x2 and 3x3 (row-major): the inverse is written only when |det| > thr.
template <class T>
bool inverse_with_check2(const T m[4], T out[4], T thr) {
  const T det = m[0] * m[3] - m[1] * m[2];
  if (!(std::fabs(det) > thr)) return false;
  out[0] = m[3] / det;
  out[1] = -m[1] / det;
  out[2] = -m[2] / det;
  out[3] = m[0] / det;
  return true;
}
template <class T>
bool inverse_with_check3(const T m[9], T out[9], T thr) {
  const T c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
  const T det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  if (!(std::fabs(det) > thr)) return false;
  const T c10 = m[2] * m[7] - m[1] * m[8], c11 = m[0] * m[8] - m[2] * m[6], c12 = m[1] * m[6] - m[0] * m[7];
  const T c20 = m[1] * m[5] - m[2] * m[4], c21 = m[2] * m[3] - m[0] * m[5], c22 = m[0] * m[4] - m[1] * m[3];
  const T cof[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i * 3 + j] = cof[j * 3 + i] / det;  // adjugate / det
  return true;
}

// StatisticalBase (common/statistic.h:80-106): running mean and n * variance.
template <class T>
struct RunStat {
  T mean = 0, nvar = 0;
  int count = 0;
  void add(T v) {
    ++count;
    const T d = v - mean;
    mean += d / T(count);
    nvar += d * d * T(count - 1) / T(count);
  }
  T variance() const { return nvar / T(count); }
};

struct OrcMonoProblem {
  int32_t num_frames, num_tracks, num_obs;
  const int32_t* constrained;  // [num_frames]
  const float* frame_poses;    // [num_frames][16] column-major world_from_camera (the addFrame argument)
  const float* points;         // [num_tracks][3]
  const int32_t* obs_frame;    // [num_obs]
  const int32_t* obs_track;    // [num_obs]
  const float* obs_xy;         // [num_obs][2]
  int32_t max_iterations;
  float min_residual, min_residual_speed;
};

struct OrcMonoResult {
  double* frame_poses;  // out [num_frames][16] world_from_camera
  double* points;       // out [num_tracks][3]
  double covariance[36];
  double residual, residual_initial, lambda;
  int32_t iterations, failed, feasible, pad;
};

template <class T>
struct MonoSolver {
  struct Frame {
    bool constrained;
    T a[6], aD[6], da[6], errA[6], eps[6], Uo[36], Ua[36];
  };
  struct Track {
    bool skip = false, removed = false;
    T b[3], bD[3], db[3], errB[3], V[9], invVa[9], info[4];
    std::vector<int> frames, obs;  // frames observing it in frame order, and the observation of each
  };
  struct Obs {
    T coord[2], err[2], w, A[12], B[6], W[18], Y[18];
  };
  int max_iterations;
  T min_residual, min_residual_speed;
  T residual = -1, residual_initial = 0, lambda = T(0.001f);
  int iteration = -1;
  bool failed = false;
  std::vector<Frame> frames;
  std::vector<Track> tracks;
  std::vector<Obs> obs;
  T prep[36] = {}, cov[36] = {};

  void load(const OrcMonoProblem* p) {
    max_iterations = p->max_iterations;
    min_residual = static_cast<T>(p->min_residual);
    min_residual_speed = static_cast<T>(p->min_residual_speed);
    frames.resize(p->num_frames);
    for (int f = 0; f < p->num_frames; ++f) {
      Frame& F = frames[f];
      F.constrained = p->constrained[f] != 0;
      const Pose<T> wc = load_pose<T>(p->frame_poses + 16 * f);
      Pose<T> cw;  // the inverse: a maps world coordinates to frame coordinates (mono_sba_solver.cpp:45)
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) cw.R[i][j] = wc.R[j][i];
        cw.t[i] = -(wc.R[0][i] * wc.t[0] + wc.R[1][i] * wc.t[1] + wc.R[2][i] * wc.t[2]);
      }
      twist_log<T>(cw, F.a);
      std::fill(F.da, F.da + 6, T(0));
      std::fill(F.errA, F.errA + 6, T(0));
      if (F.constrained) std::copy(F.a, F.a + 6, F.aD);
    }
    tracks.resize(p->num_tracks);
    for (int t = 0; t < p->num_tracks; ++t) {
      Track& K = tracks[t];
      for (int d = 0; d < 3; ++d) K.b[d] = static_cast<T>(p->points[3 * t + d]);
      std::fill(K.db, K.db + 3, T(0));
      K.info[0] = K.info[3] = T(1);
      K.info[1] = K.info[2] = T(0);
    }
    obs.resize(p->num_obs);
    // frames are added in order, each listing its tracks: a track's frames are in frame order
    std::vector<int> order(p->num_obs);
    for (int o = 0; o < p->num_obs; ++o) order[o] = o;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return p->obs_frame[x] < p->obs_frame[y]; });
    for (int o : order) {
      Obs& O = obs[o];
      O.coord[0] = static_cast<T>(p->obs_xy[2 * o]);
      O.coord[1] = static_cast<T>(p->obs_xy[2 * o + 1]);
      O.err[0] = O.err[1] = T(0);
      O.w = T(1);
      tracks[p->obs_track[o]].frames.push_back(p->obs_frame[o]);
      tracks[p->obs_track[o]].obs.push_back(o);
    }
  }

  // project world point b through the camera with twist a (camera_from_world)
  static void project(const T a[6], const T b[3], T out[2]) {
    const Pose<T> M = twist_exp<T>(a);
    T q[3];
    apply(M, b, q);
    const T z = avoid_zero(q[2]);
    out[0] = q[0] / z;
    out[1] = q[1] / z;
  }

  void filter_hither() {  // mono_sba_solver.cpp:346-361
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (int f : K.frames) {
        const Pose<T> M = twist_exp<T>(frames[f].a);
        T q[3];
        apply(M, K.b, q);
        if (!(q[2] > T(kHither))) {
          K.skip = true;
          break;
        }
      }
    }
  }

  void update_aDbD() {  // mono_sba_solver.cpp:770-786
    for (Frame& F : frames)
      if (!F.constrained) twist_add<T>(F.a, F.da, F.aD);
    for (Track& K : tracks)
      if (!K.skip)
        for (int d = 0; d < 3; ++d) K.bD[d] = K.b[d] + K.db[d];
  }

  void update() {  // mono_sba_solver.cpp:788-810
    for (Frame& F : frames)
      if (!F.constrained) std::copy(F.aD, F.aD + 6, F.a);
    for (Track& K : tracks) {
      if (K.removed) continue;
      if (K.skip)
        K.skip = false;
      else
        std::copy(K.bD, K.bD + 3, K.b);
    }
  }

  static T huber(T s) { return loss_rho<T>(1, s, T(kMonoHuber)); }
  static T dhuber(T s) { return loss_drho<T>(1, s, T(kMonoHuber)); }
  static T quad(const T e[2], const T L[4]) { return e[0] * (L[0] * e[0] + L[1] * e[1]) + e[1] * (L[2] * e[0] + L[3] * e[1]); }

  T calc_residual() {  // mono_sba_solver.cpp:743-768
    int total = 0;
    T r = 0;
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        Obs& O = obs[K.obs[k]];
        T proj[2];
        project(frames[K.frames[k]].aD, K.bD, proj);
        O.err[0] = O.coord[0] - proj[0];
        O.err[1] = O.coord[1] - proj[1];
        r += huber(quad(O.err, K.info));
        ++total;
      }
    }
    return total != 0 ? r / T(total) : T(std::numeric_limits<float>::max());
  }

  // mono_sba_solver.cpp:363-432: the covariance of each track's residuals over its frames (variances floored at
  // 1e-5; with more than three samples the cross term, then an LDLT with the floor on D, reassembled without the
  // pivot permutation as the reference does), inverted when its determinant exceeds FLT_EPSILON, else the diagonal
  void calc_information() {
    const T thresh = T(1e-5f);
    for (Track& K : tracks) {
      if (K.skip) continue;
      RunStat<T> sx, sy, sxy;
      for (int o : K.obs) {
        sx.add(obs[o].err[0]);
        sy.add(obs[o].err[1]);
        sxy.add(obs[o].err[0] * obs[o].err[1]);
      }
      T c[4] = {std::max(sx.variance(), thresh), 0, 0, std::max(sy.variance(), thresh)};
      c[0] = sx.variance() < thresh ? thresh : sx.variance();
      c[3] = sy.variance() < thresh ? thresh : sy.variance();
      if (sx.count > 3) {
        c[1] = c[2] = sxy.mean - sx.mean * sy.mean;
        // Eigen::LDLT with pivoting: the larger diagonal first (the first on a tie)
        const bool swap = std::fabs(c[3]) > std::fabs(c[0]);
        const T p00 = swap ? c[3] : c[0], p11 = swap ? c[0] : c[3], p01 = c[1];
        const T d0 = p00, l = p01 / p00, d1 = p11 - l * p01;
        const T e0 = std::max(d0, thresh), e1 = std::max(d1, thresh);
        c[0] = e0;
        c[1] = c[2] = l * e0;
        c[3] = l * l * e0 + e1;
      }
      if (!inverse_with_check2<T>(c, K.info, T(kFltEps))) {
        K.info[0] = T(1) / c[0];
        K.info[3] = T(1) / c[3];
        K.info[1] = K.info[2] = T(0);
      }
    }
  }

  void compute_weights() {  // mono_sba_solver.cpp:435-444 (every track, skipped or not)
    for (Track& K : tracks)
      for (int o : K.obs) obs[o].w = dhuber(quad(obs[o].err, K.info));
  }

  void calc_derivatives() {  // mono_sba_solver.cpp:704-741, math/twist.h:152-206
    std::vector<Pose<T>> M(frames.size());
    std::vector<Pose<T>> Rn(frames.size());
    for (size_t f = 0; f < frames.size(); ++f) {
      M[f] = twist_exp<T>(frames[f].a);
      nearest_rotation<T>(M[f].R, Rn[f].R);
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        const int f = K.frames[k];
        Obs& O = obs[K.obs[k]];
        T q[3];
        apply(M[f], K.b, q);
        const T iz = T(1) / avoid_zero(q[2]);
        const T x = q[0] * iz, y = q[1] * iz;
        for (int j = 0; j < 3; ++j) {
          O.B[j] = (Rn[f].R[0][j] - Rn[f].R[2][j] * x) * iz;
          O.B[3 + j] = (Rn[f].R[1][j] - Rn[f].R[2][j] * y) * iz;
        }
        if (!frames[f].constrained) {
          // -am: the reference's camera derivatives with the sign of HZ's convention
          const T am[12] = {x * y, -x * x - T(1), y, -iz, T(0), x * iz, y * y + T(1), -(x * y), -x, T(0), -iz, y * iz};
          for (int e = 0; e < 12; ++e) O.A[e] = -am[e];
        }
      }
    }
  }

  // tmp = A^T (w info) (6x2), then accumulations (mono_sba_solver.cpp:586-678)
  void calc_UVW() {
    for (Frame& F : frames) {
      if (F.constrained) continue;
      std::fill(F.Uo, F.Uo + 36, T(0));
      std::fill(F.errA, F.errA + 6, T(0));
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        Frame& F = frames[K.frames[k]];
        if (F.constrained) continue;
        const Obs& O = obs[K.obs[k]];
        T tmp[6][2];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 2; ++c) tmp[r][c] = O.A[r] * (O.w * K.info[c]) + O.A[6 + r] * (O.w * K.info[2 + c]);
        for (int r = 0; r < 6; ++r) {
          for (int c = 0; c < 6; ++c) F.Uo[r * 6 + c] += tmp[r][0] * O.A[c] + tmp[r][1] * O.A[6 + c];
          F.errA[r] += tmp[r][0] * O.err[0] + tmp[r][1] * O.err[1];
        }
      }
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      std::fill(K.V, K.V + 9, T(0));
      std::fill(K.errB, K.errB + 3, T(0));
      for (int o : K.obs) {
        const Obs& O = obs[o];
        T tmp[3][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 2; ++c) tmp[r][c] = O.B[r] * (O.w * K.info[c]) + O.B[3 + r] * (O.w * K.info[2 + c]);
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) K.V[r * 3 + c] += tmp[r][0] * O.B[c] + tmp[r][1] * O.B[3 + c];
          K.errB[r] += tmp[r][0] * O.err[0] + tmp[r][1] * O.err[1];
        }
      }
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        if (frames[K.frames[k]].constrained) continue;
        Obs& O = obs[K.obs[k]];
        T Li[2][2];  // info * w
        for (int a = 0; a < 2; ++a)
          for (int c = 0; c < 2; ++c) Li[a][c] = K.info[a * 2 + c] * O.w;
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c) {
            const T m0 = O.A[r] * Li[0][0] + O.A[6 + r] * Li[1][0];
            const T m1 = O.A[r] * Li[0][1] + O.A[6 + r] * Li[1][1];
            O.W[r * 3 + c] = m0 * O.B[c] + m1 * O.B[3 + c];
          }
      }
    }
  }

  void step_UV() {  // mono_sba_solver.cpp:446-510, invert3x3 :892-899
    for (Frame& F : frames) {
      if (F.constrained) continue;
      std::copy(F.Uo, F.Uo + 36, F.Ua);
      for (int i = 0; i < 6; ++i) F.Ua[i * 7] *= T(1) + lambda;
    }
    const T thr = T(1) / std::max(T(1e6f), lambda * lambda * lambda);
    for (Track& K : tracks) {
      if (K.skip) continue;
      T V[9];
      std::copy(K.V, K.V + 9, V);
      for (int i = 0; i < 3; ++i) V[i * 4] *= T(1) + lambda;
      if (!inverse_with_check3<T>(V, K.invVa, thr)) K.skip = true;
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        if (frames[K.frames[k]].constrained) continue;
        Obs& O = obs[K.obs[k]];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c)
            O.Y[r * 3 + c] = O.W[r * 3] * K.invVa[c] + O.W[r * 3 + 1] * K.invVa[3 + c] + O.W[r * 3 + 2] * K.invVa[6 + c];
      }
    }
  }

  // JacobiSVD solve with threshold n FLT_EPSILON (IterativeSolver, common/types.h:250-290) of a symmetric system:
  // false when a singular value is at or below the threshold (rank < n); else x with up to 5 refinement steps
  // kept while the residual norm decreases
  static bool svd_solve(int n, const std::vector<T>& S, const std::vector<T>& bvec, std::vector<T>& x) {
    std::vector<T> a(S), ev(n), q(static_cast<size_t>(n) * n);
    sym_jacobi<T>(n, a.data(), ev.data(), q.data());
    T smax = 0;
    for (int k = 0; k < n; ++k) smax = std::max(smax, std::fabs(ev[k]));
    const T thr = T(n) * T(kFltEps) * smax;
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += std::fabs(ev[k]) > thr;
    if (rank < n) return false;
    auto psolve = [&](const std::vector<T>& rhs, std::vector<T>& out) {
      out.assign(n, T(0));
      for (int k = 0; k < n; ++k) {
        T s = 0;
        for (int i = 0; i < n; ++i) s += q[i * n + k] * rhs[i];
        s /= ev[k];
        for (int i = 0; i < n; ++i) out[i] += q[i * n + k] * s;
      }
    };
    auto norm = [](const std::vector<T>& v) {
      T s = 0;
      for (T e : v) s += e * e;
      return std::sqrt(s);
    };
    std::vector<T> xt(n, T(0)), bd(bvec), d;
    x.assign(n, T(0));
    T min_norm = std::numeric_limits<T>::max(), nrm = norm(bvec);
    for (int it = 0; nrm < min_norm && it < 5; ++it) {
      min_norm = nrm;
      x = xt;
      psolve(bd, d);
      for (int i = 0; i < n; ++i) xt[i] += d[i];
      for (int i = 0; i < n; ++i) {
        T s = bvec[i];
        for (int j = 0; j < n; ++j) s -= S[static_cast<size_t>(i) * n + j] * xt[j];
        bd[i] = s;
      }
      nrm = norm(bd);
      if (nrm < T(kFltEps)) {
        x = xt;
        break;
      }
    }
    return true;
  }

  bool calc_adjustments() {  // mono_sba_solver.cpp:812-890 with calculateEps :681-702 and calculateSMatrix :812-870
    std::vector<int> idx(frames.size(), -1);
    int nfree = 0;
    for (size_t f = 0; f < frames.size(); ++f) {
      Frame& F = frames[f];
      if (F.constrained) continue;
      std::copy(F.errA, F.errA + 6, F.eps);
      idx[f] = nfree++;
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t k = 0; k < K.frames.size(); ++k) {
        Frame& F = frames[K.frames[k]];
        if (F.constrained) continue;
        const Obs& O = obs[K.obs[k]];
        for (int r = 0; r < 6; ++r) F.eps[r] -= O.Y[r * 3] * K.errB[0] + O.Y[r * 3 + 1] * K.errB[1] + O.Y[r * 3 + 2] * K.errB[2];
      }
    }
    const int n = 6 * nfree;
    if (n == 0) {
      for (Track& K : tracks)
        if (!K.removed) std::fill(K.db, K.db + 3, T(0));
      return false;
    }
    std::vector<T> S(static_cast<size_t>(n) * n, T(0)), b(n);
    for (size_t f = 0; f < frames.size(); ++f) {
      if (idx[f] < 0) continue;
      for (int r = 0; r < 6; ++r) {
        b[6 * idx[f] + r] = frames[f].eps[r];
        for (int c = 0; c < 6; ++c) S[static_cast<size_t>(6 * idx[f] + r) * n + 6 * idx[f] + c] = frames[f].Ua[r * 6 + c];
      }
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      for (size_t j = 0; j < K.frames.size(); ++j) {
        if (idx[K.frames[j]] < 0) continue;
        const Obs& Oj = obs[K.obs[j]];
        for (size_t k = 0; k < K.frames.size(); ++k) {
          if (idx[K.frames[k]] < 0) continue;
          const Obs& Ok = obs[K.obs[k]];
          for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
              S[static_cast<size_t>(6 * idx[K.frames[j]] + r) * n + 6 * idx[K.frames[k]] + c] -=
                  Oj.Y[r * 3] * Ok.W[c * 3] + Oj.Y[r * 3 + 1] * Ok.W[c * 3 + 1] + Oj.Y[r * 3 + 2] * Ok.W[c * 3 + 2];
        }
      }
    }
    const int jj = n - 6;
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) prep[r * 6 + c] = S[static_cast<size_t>(jj + r) * n + jj + c];
    for (int r = 0; r < 6; ++r) prep[r * 7] /= T(1) + lambda;
    std::vector<T> x;
    if (!svd_solve(n, S, b, x)) return false;
    for (size_t f = 0; f < frames.size(); ++f) {
      Frame& F = frames[f];
      if (F.constrained) {
        std::fill(F.da, F.da + 6, T(0));
        continue;
      }
      for (int i = 0; i < 6; ++i) F.da[i] = x[6 * idx[f] + i];
    }
    for (Track& K : tracks) {
      if (K.skip) continue;
      T wda[3] = {0, 0, 0};
      for (size_t k = 0; k < K.frames.size(); ++k) {
        const Frame& F = frames[K.frames[k]];
        if (F.constrained) continue;
        const Obs& O = obs[K.obs[k]];
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 6; ++r) wda[c] += O.W[r * 3 + c] * F.da[r];
      }
      const T pt[3] = {K.errB[0] - wda[0], K.errB[1] - wda[1], K.errB[2] - wda[2]};
      for (int r = 0; r < 3; ++r) K.db[r] = K.invVa[r * 3] * pt[0] + K.invVa[r * 3 + 1] * pt[1] + K.invVa[r * 3 + 2] * pt[2];
    }
    return true;
  }

  void covariance_from_prep() {  // CompleteOrthogonalDecomposition(prep).pseudoInverse()
    std::vector<T> P(prep, prep + 36);
    for (int r = 0; r < 6; ++r)  // the reduced system is symmetric: take its symmetric part
      for (int c = r + 1; c < 6; ++c) P[r * 6 + c] = P[c * 6 + r] = (P[r * 6 + c] + P[c * 6 + r]) / T(2);
    sym_pinv<T>(6, P.data(), T(6) * T(kFltEps), cov);
  }

  int solve(bool& feasible) {  // mono_sba_solver.cpp:205-301
    int count = max_iterations;
    feasible = true;
    filter_hither();
    update_aDbD();
    calc_residual();
    calc_information();
    compute_weights();
    residual = calc_residual();
    const T fmax = T(std::numeric_limits<float>::max());
    if (residual == fmax) {
      for (int i = 0; i < 36; ++i) cov[i] = T(i % 7 == 0);
      feasible = false;
      return 0;
    }
    residual_initial = residual;
    bool new_values = true;
    T old = residual;
    while (min_residual <= residual && count >= 0 && lambda <= T(kMonoLambdaMax)) {
      if (new_values) {
        calc_derivatives();
        calc_UVW();
      }
      step_UV();
      if (calc_adjustments()) {
        update_aDbD();
        residual = calc_residual();
        if (residual == fmax) {
          feasible = false;
          return 0;
        }
      } else {
        if (residual == fmax) {
          feasible = false;
          return 0;
        }
        residual *= T(2);
      }
      const T delta = std::fabs(residual - old);
      if (residual > old) {
        lambda *= T(10);
        new_values = false;
      } else {
        if (lambda > T(1e-25f)) lambda /= T(10);
        old = residual;
        new_values = true;
        update();
        filter_hither();
        calc_information();
        compute_weights();
        covariance_from_prep();
      }
      if (min_residual_speed * residual > delta) break;
      --count;
    }
    residual = old;
    failed = lambda > T(kMonoLambdaMax) || count < 0;
    if (failed) covariance_from_prep();
    return iteration = max_iterations - count;
  }

  void export_state(const OrcMonoProblem* p, OrcMonoResult* r) const {
    for (int f = 0; f < p->num_frames && r->frame_poses; ++f) {
      const Pose<T> cw = twist_exp<T>(frames[f].a);
      Pose<T> wc;
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) wc.R[i][j] = cw.R[j][i];
        wc.t[i] = -(cw.R[0][i] * cw.t[0] + cw.R[1][i] * cw.t[1] + cw.R[2][i] * cw.t[2]);
      }
      store_pose(wc, r->frame_poses + 16 * f);
    }
    for (int t = 0; t < p->num_tracks && r->points; ++t)
      for (int d = 0; d < 3; ++d) r->points[3 * t + d] = static_cast<double>(tracks[t].b[d]);
    for (int i = 0; i < 36; ++i) r->covariance[i] = static_cast<double>(cov[i]);
    r->residual = static_cast<double>(residual);
    r->residual_initial = static_cast<double>(residual_initial);
    r->lambda = static_cast<double>(lambda);
    r->failed = failed;
  }
};

template <class T>
int mono_impl(const OrcMonoProblem* p, OrcMonoResult* r) {
  MonoSolver<T> s;
  s.load(p);
  bool feasible = true;
  const int it = s.solve(feasible);
  r->iterations = it;
  r->feasible = feasible;
  s.export_state(p, r);
  return feasible ? 0 : 2;
}

}  // namespace

extern "C" {

int orc_solve_f64(const OrcProblem* p, const OrcOptions* o, OrcResult* r) { return solve_impl<double>(p, o, r); }
// The FP64 solve from an FP64 initial state (points [np][3], poses [nc][16] column-major).
int orc_solve_state_f64(const OrcProblem* p, const OrcOptions* o, const double* sp, const double* sq,
                        const double* oxy, const double* oinfo, OrcResult* r) {
  return solve_impl<double>(p, o, r, sp, sq, oxy, oinfo);
}
int orc_solve_f32(const OrcProblem* p, const OrcOptions* o, OrcResult* r) { return solve_impl<float>(p, o, r); }

// Per-observation residual (2), pose Jacobian (2x6 row-major), point Jacobian (2x3), weight, at the input state.
int orc_linearize_f64(const OrcProblem* p, const OrcOptions* o, double* r, double* Jc, double* Jp, double* w) {
  Solver<double> s;
  s.load(p, o);
  Model<double> m;
  s.update_model(m);
  std::copy(m.r.begin(), m.r.end(), r);
  std::copy(m.Jc.begin(), m.Jc.end(), Jc);
  std::copy(m.Jp.begin(), m.Jp.end(), Jp);
  std::copy(m.w.begin(), m.w.end(), w);
  return 0;
}

// One LM step at the input state with damping lambda (the phases of cpu.cpp:479-494):
// S [n*n] damped reduced system, e [n], dc [n], dp [3 npf];
// scalars [9]: trial cost, skipped, delta^T H delta, delta^T diag(H) delta, max|dp|, max|dc|, current cost, ok,
// predicted relative reduction (ComputePredictedRelativeReduction, cpu.cpp:385-433, at the current cost).
int orc_step_f64(const OrcProblem* p, const OrcOptions* o, double lambda, double* S, double* e, double* dc,
                 double* dp, double* scalars) {
  Solver<double> s;
  s.load(p, o);
  Model<double> m;
  s.update_model(m);
  Full<double> F;
  s.build_full(F, m);
  Solver<double>::Reduced R;
  s.build_reduced(R, F, lambda);
  if (S) std::copy(R.S.begin(), R.S.end(), S);
  if (e) std::copy(R.e.begin(), R.e.end(), e);
  Solver<double>::Update u;
  s.compute_update(u, R);
  if (dc) std::copy(u.dc.begin(), u.dc.end(), dc);
  if (dp) std::copy(u.dp.begin(), u.dp.end(), dp);
  int skipped = 0;
  const double current = s.evaluate_cost(nullptr, nullptr);
  const double trial = s.evaluate_cost(&u, &skipped);
  double hess = 0, scal = 0;
  const double pred = s.predicted(current > 0 ? current : 1.0, lambda, u, F, &hess, &scal);
  if (scalars) {
    scalars[0] = trial;
    scalars[1] = skipped;
    scalars[2] = hess;
    scalars[3] = scal;
    scalars[4] = Solver<double>::inf_norm(u.dp);
    scalars[5] = Solver<double>::inf_norm(u.dc);
    scalars[6] = current;
    scalars[7] = u.ok ? 1.0 : 0.0;
    scalars[8] = pred;
  }
  return u.ok ? 0 : 2;
}

// The predicted relative reduction of the step at the input state with damping lambda, divided by a given
// current cost (cpu.cpp:385-433 and sba.cpp:688-727 divide by the caller's current cost; a vanished cost gives
// a value that is not finite, which sba_test.cpp:1274-1317 pins).
double orc_predicted_f64(const OrcProblem* p, const OrcOptions* o, double lambda, double current) {
  Solver<double> s;
  s.load(p, o);
  Model<double> m;
  s.update_model(m);
  Full<double> F;
  s.build_full(F, m);
  Solver<double>::Reduced R;
  s.build_reduced(R, F, lambda);
  Solver<double>::Update u;
  s.compute_update(u, R);
  return s.predicted(current, lambda, u, F, nullptr, nullptr);
}

// Cost at the input state moved by (dc, dp) (either may be null for a zero step).
double orc_cost_f64(const OrcProblem* p, const OrcOptions* o, const double* dc, const double* dp, int32_t* skipped) {
  Solver<double> s;
  s.load(p, o);
  Solver<double>::Update u;
  u.dc.assign(6 * static_cast<size_t>(s.ncf), 0.0);
  u.dp.assign(3 * static_cast<size_t>(s.npf), 0.0);
  if (dc) std::copy(dc, dc + u.dc.size(), u.dc.begin());
  if (dp) std::copy(dp, dp + u.dp.size(), u.dp.begin());
  u.pose.resize(s.ncf);
  for (int j = 0; j < s.ncf; ++j) u.pose[j] = twist_exp<double>(&u.dc[6 * j]);
  int sk = 0;
  const double c = s.evaluate_cost(&u, &sk);
  if (skipped) *skipped = sk;
  return c;
}

int orc_chol_solve_f64(int32_t n, const double* A, const double* b, double* x) { return chol_solve<double>(n, A, b, x) ? 0 : 1; }
int orc_chol_solve_f32(int32_t n, const float* A, const float* b, float* x) { return chol_solve<float>(n, A, b, x) ? 0 : 1; }
int orc_sym_pinv_f64(int32_t n, const double* A, double thr, double* out) { return sym_pinv<double>(n, A, thr, out); }
int orc_point_pinv_f64(const double* A, double* out) { return point_pinv<double>(A, out); }

void orc_exp_f64(const double* tw, double* m16) { store_pose(twist_exp<double>(tw), m16); }

void orc_nearest_rotation_f64(const double* M9, double* R9) {
  double M[3][3], R[3][3];
  for (int i = 0; i < 9; ++i) M[i / 3][i % 3] = M9[i];
  nearest_rotation<double>(M, R);
  for (int i = 0; i < 9; ++i) R9[i] = R[i / 3][i % 3];
}

// Implicit gradient at the input state (which should be a converged solution): g_tangent (6 per free key
// frame, then 3 per free point) -> adjoint a, dL/dxy (2 per observation), dL/dinfo (3 per observation).
// Returns the number of conjugate-gradient iterations (negative when the preconditioner failed).
int orc_implicit_grad_f64(const OrcProblem* p, const OrcOptions* o, const double* sp, const double* sq,
                          const double* oxy, const double* oinfo, const double* g, double* a, double* dxy,
                          double* dinfo, int32_t max_it, double tol) {
  Solver<double> s;
  s.load(p, o);
  s.override_state(sp, sq);
  s.override_observations(oxy, oinfo);
  Adjoint adj(s);
  const int it = adj.pcg(g, a, max_it, tol);
  if (it < 0) return it;
  adj.observation_grads(a, dxy, dinfo);
  return it;
}

// y = H x with the full Hessian of the cost at the input state (tangent coordinates as above).
void orc_hvp_f64(const OrcProblem* p, const OrcOptions* o, const double* sp, const double* sq, const double* x, double* y) {
  Solver<double> s;
  s.load(p, o);
  s.override_state(sp, sq);
  Adjoint adj(s);
  adj.hvp(x, y);
}

// Gradient of the cost (tangent coordinates), -2 sum rho' J^T L r.
void orc_grad_f64(const OrcProblem* p, const OrcOptions* o, const double* sp, const double* sq, const double* oxy,
                  const double* oinfo, double* g) {
  Solver<double> s;
  s.load(p, o);
  s.override_state(sp, sq);
  s.override_observations(oxy, oinfo);
  Adjoint adj(s);
  std::fill(g, g + adj.nx, 0.0);
  for (const ObsLin& L : adj.obs) {
    if (!L.valid) continue;
    const double Lr[2] = {L.L[0][0] * L.r[0] + L.L[0][1] * L.r[1], L.L[1][0] * L.r[0] + L.L[1][1] * L.r[1]};
    double h[9];
    for (int k = 0; k < 9; ++k) h[k] = -2 * L.w * (L.J[0][k] * Lr[0] + L.J[1][k] * Lr[1]);
    adj.scatter(L, h, g);
  }
}

double orc_loss_f64(int32_t kind, double s, double d, int32_t derivative) {
  return derivative ? loss_drho<double>(kind, s, d) : loss_rho<double>(kind, s, d);
}

// Landmark winsorizing (pipelines/winsorizer.cpp:23-73): squared reprojection errors, in normalized
// coordinates, of the observations made from key frame kf at the state (points [np][3], poses [nc][16]
// column-major; the problem's own when null), skipping points inside the near plane (IsDepthInFront,
// epipolar/near_plane.h:38). With at least 50 of them (winsorizer.cpp:62-64), the int(5 n / 100) largest of
// the sorted (error, point, observation) triples reset their points when above 0.0025^2
// (winsorizer.cpp:66-73; the reference orders equal errors by landmark address). reset [np] 0/1; returns the
// number of reset points.
int orc_winsorize_f64(const OrcProblem* p, int32_t kf, const double* pts, const double* poses, uint8_t* reset) {
  std::fill(reset, reset + p->num_points, uint8_t(0));
  double M[16];
  for (int k = 0; k < 16; ++k) M[k] = poses ? poses[16 * kf + k] : static_cast<double>(p->rig_from_world[16 * kf + k]);
  std::vector<std::tuple<double, int, int>> errs;
  for (int o = 0; o < p->num_observations; ++o) {
    if (p->pose_ids[o] != kf) continue;
    const int i = p->point_ids[o];
    double x[3];
    for (int a = 0; a < 3; ++a) x[a] = pts ? pts[3 * i + a] : static_cast<double>(p->points[3 * i + a]);
    double r[3], c[3];
    for (int a = 0; a < 3; ++a) r[a] = M[a] * x[0] + M[4 + a] * x[1] + M[8 + a] * x[2] + M[12 + a];
    const float* Cm = p->camera_from_rig + 16 * p->camera_ids[o];
    for (int a = 0; a < 3; ++a)
      c[a] = static_cast<double>(Cm[a]) * r[0] + static_cast<double>(Cm[4 + a]) * r[1] +
             static_cast<double>(Cm[8 + a]) * r[2] + static_cast<double>(Cm[12 + a]);
    if (!(c[2] > static_cast<double>(kHither))) continue;
    const double u = c[0] / c[2] - p->observation_xys[2 * o], v = c[1] / c[2] - p->observation_xys[2 * o + 1];
    errs.emplace_back(u * u + v * v, i, o);
  }
  const int n = static_cast<int>(errs.size());
  if (n < 50) return 0;
  std::sort(errs.begin(), errs.end());
  const int cut = static_cast<int>(5.f * static_cast<float>(n) / 100.f);
  const double min2 = static_cast<double>(0.0025f * 0.0025f);
  for (int j = 0; j < cut; ++j) {
    const auto& t = errs[n - 1 - j];
    if (std::get<0>(t) > min2) reset[std::get<1>(t)] = 1;
  }
  int count = 0;
  for (int i = 0; i < p->num_points; ++i) count += reset[i];
  return count;
}

// The mono SBA solver (libs/sba/mono_sba_solver.cpp) on one window: returns 0, or 2 when no observation is usable.
int orc_mono_solve_f64(const OrcMonoProblem* p, OrcMonoResult* r) { return mono_impl<double>(p, r); }
int orc_mono_solve_f32(const OrcMonoProblem* p, OrcMonoResult* r) { return mono_impl<float>(p, r); }

// Pieces of the mono solver for the pins: the projection derivatives (2x6 camera, 2x3 point, row-major) of world
// point b seen through the camera with twist a; the twist Log of a pose (column-major 4x4); the twist sum.
void orc_mono_derivs_f64(const double* a, const double* b, double* A, double* B) {
  MonoSolver<double> s;
  s.frames.resize(1);
  s.frames[0].constrained = false;
  std::copy(a, a + 6, s.frames[0].a);
  s.tracks.resize(1);
  std::copy(b, b + 3, s.tracks[0].b);
  s.tracks[0].frames = {0};
  s.tracks[0].obs = {0};
  s.obs.resize(1);
  s.calc_derivatives();
  std::copy(s.obs[0].A, s.obs[0].A + 12, A);
  std::copy(s.obs[0].B, s.obs[0].B + 6, B);
}
void orc_twist_log_f64(const double* m16, double* tw) {
  Pose<double> P;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) P.R[r][c] = m16[c * 4 + r];
    P.t[r] = m16[12 + r];
  }
  twist_log<double>(P, tw);
}
void orc_twist_add_f64(const double* lhs, const double* rhs, double* out) { twist_add<double>(lhs, rhs, out); }

}  // extern "C"
