// The mono SBA solve of one window (libs/sba/mono_sba_solver.cpp:205-301; Hartley & Zisserman 2nd ed. Algorithm
// A6.4), written once for one workgroup: every phase is a loop over frames, tracks or observations strided by the
// workgroup's threads, with a barrier between phases. On the GPU (sba_mono.hip) a workgroup of kMonoThreads runs
// it per window; compiled as host code with one "thread" it runs the same arithmetic sequentially (the CPU check of
// tests/cpp/mono_host_check.cpp). FP64 throughout.
//
// Frames carry the twist a of camera_from_world; updates are a <- Log(Exp(da) Exp(a)) (math/twist.h:67-73). Tracks
// carry a world point and a 2x2 information matrix re-estimated from their residuals; the loss is Huber with
// threshold 0.5; damping multiplies the U and V diagonals by (1 + lambda) and lambda moves x10 / 10; the reduced
// system is solved through its eigen-decomposition with the rank threshold n FLT_EPSILON and up to five refinement
// steps (IterativeSolver, common/types.h:250-290); the covariance is the pseudo-inverse of the last free frame's
// undamped block of the reduced system.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define MONO_SYNC() __syncthreads()
#else
#define MONO_SYNC() ((void)0)
#endif

namespace sba_mono {

constexpr int kMonoThreads = 256;
constexpr int kPart = 42;  // partial sums per thread: a frame's U (36) and e (6), or a reduced-system block (36)
constexpr double kFltEps = 1.1920928955078125e-07;     // FLT_EPSILON, common/types.h:203-206
constexpr double kFltMax = 3.4028234663852886e+38;     // std::numeric_limits<float>::max()
constexpr double kHither = 0.1;                        // epipolar/near_plane.h:33-38
constexpr double kLambdaMax = 1.0e+7;                  // mono_sba_solver.cpp:26 (as float 1e7)
constexpr double kHuber = 0.5;                         // mono_sba_solver.cpp:214
constexpr double kLambdaMin = 1.0000000180025095e-25;  // 1e-25f

// One window's buffers (device or host memory). Observations are referenced through two orders: by track (the
// track's frames in frame order) and by frame (its tracks in track order); `blk` lists, for every ordered pair of
// free frames (j, k), the (observation in j, observation in k) pairs of the tracks both observe, in track order.
struct MonoWin {
  int32_t nF, nT, nO, nFree, max_iterations;
  double min_residual, min_residual_speed;
  // frames
  const int32_t* constrained;  // [nF]
  const int32_t* free_idx;     // [nF] index among the free frames, -1 when constrained
  const float* pose_in;        // [nF][16] world_from_camera, column-major
  double* fa;                  // [nF][6] a
  double* faD;                 // [nF][6]
  double* fda;                 // [nF][6]
  double* ferrA;               // [nF][6]
  double* feps;                // [nF][6]
  double* fUo;                 // [nF][36]
  double* fUa;                 // [nF][36]
  double* fM;                  // [nF][12] Exp(a): R row-major, t
  double* fRn;                 // [nF][9] nearest rotation of R
  const int32_t* frame_obs;    // [nF+1] into fobs
  const int32_t* fobs;         // [nO] observations of each frame, in track order
  // tracks
  const float* pts_in;         // [nT][3]
  const int32_t* track_obs;    // [nT+1] into tobs
  const int32_t* tobs;         // [nO] observations of each track, in frame order
  int32_t* tskip;              // [nT]
  double* tb;                  // [nT][3]
  double* tbD;                 // [nT][3]
  double* tdb;                 // [nT][3]
  double* terrB;               // [nT][3]
  double* tV;                  // [nT][9]
  double* tinvVa;              // [nT][9]
  double* tinfo;               // [nT][4]
  // observations
  const int32_t* oframe;       // [nO]
  const int32_t* otrack;       // [nO]
  const float* oxy;            // [nO][2]
  double* oerr;                // [nO][2]
  double* ow;                  // [nO]
  double* oA;                  // [nO][12]
  double* oB;                  // [nO][6]
  double* oW;                  // [nO][18]
  double* oY;                  // [nO][18]
  // reduced system
  const int32_t* blk;          // [nFree*nFree + 1] into bpair
  const int2* bpair;           // (observation in j, observation in k)
  double* S;                   // [n][n]
  double* Sb;                  // [n]
  double* Q;                   // [n][n] eigenvectors
  double* E;                   // [n][n] work copy of S
  double* E2;                  // [n][n] the Jacobi sweeps' second buffer (when E does not fit in LDS)
  double* ev;                  // [n]
  double* x;                   // [n] and four more work vectors
  double* xt;
  double* bd;
  double* dv;
  double* rot;                 // [2 (n + 1)] the Jacobi rotations of one step (p, r, c, s), also psolve's scratch
  double* part;                // [kMonoThreads * kPart] per-thread partial sums of the sliced loops
  // outputs
  double* prep;                // [36] the last free frame's block of the reduced system (undamped diagonal)
  double* cov;                 // [36]
  double* out_scalars;         // [8]: residual, residual_initial, lambda, iterations, failed, feasible
};

// One window of a batched launch: its buffers and where its solved poses and points go.
struct MonoJob {
  MonoWin w;
  float* poses_out;  // [nF][16] world_from_camera
  float* pts_out;    // [nT][3]
};

// ---- small FP64 helpers -------------------------------------------------------------------------------------------

__host__ __device__ inline double avoid_zero(double a) {  // common/types.h:219-221
  return fabs(a) > kFltEps ? a : copysign(kFltEps, a);
}
__host__ __device__ inline double huber(double s) {  // robust_cost_function.cpp:62-74, delta 0.5
  return s < kHuber * kHuber ? 0.5 * s : kHuber * sqrt(s) - 0.5 * kHuber * kHuber;
}
__host__ __device__ inline double dhuber(double s) {  // robust_cost_function.cpp:76-87
  return s < kHuber * kHuber ? 0.5 : 0.5 * kHuber / sqrt(s);
}
__host__ __device__ inline double quad(const double* e, const double* L) {
  return e[0] * (L[0] * e[0] + L[1] * e[1]) + e[1] * (L[2] * e[0] + L[3] * e[1]);
}

// Exp of a twist (w, v) (math/twist.cpp:23-41, theta clamped at FLT_EPSILON): R (row-major) and t.
__host__ __device__ inline void twist_exp(const double* tw, double* R, double* t) {
  const double* w = tw;
  const double* v = tw + 3;
  double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  th = th > kFltEps ? th : kFltEps;
  const double c = cos(th), st = sin(th) / th;
  const double n[3] = {w[0] / th, w[1] / th, w[2] / th};
  const double ws[3] = {w[0] * st, w[1] * st, w[2] * st};
  const double K[9] = {c, -ws[2], ws[1], ws[2], c, -ws[0], -ws[1], ws[0], c};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[3 * i + j] = K[3 * i + j] + (1.0 - c) * n[i] * n[j];
  const double nxv[3] = {n[1] * v[2] - n[2] * v[1], n[2] * v[0] - n[0] * v[2], n[0] * v[1] - n[1] * v[0]};
  const double nxnxv[3] = {n[1] * nxv[2] - n[2] * nxv[1], n[2] * nxv[0] - n[0] * nxv[2],
                           n[0] * nxv[1] - n[1] * nxv[0]};
  for (int i = 0; i < 3; ++i) t[i] = v[i] + ((1.0 - c) / th * nxv[i] + (1.0 - st) * nxnxv[i]);
}

// Cyclic Jacobi eigen-decomposition of a small symmetric m x m matrix (Golub & Van Loan, Algorithm 8.5.3), one
// thread: a (row-major, destroyed) -> eigenvalues ev, eigenvectors the columns of q.
__host__ __device__ inline void small_jacobi(int m, double* a, double* ev, double* q) {
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) q[i * m + j] = i == j ? 1.0 : 0.0;
  double fro = 0.0;
  for (int i = 0; i < m * m; ++i) fro += a[i] * a[i];
  const double tol = 2.220446049250313e-16 * sqrt(fro);
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j) off += a[i * m + j] * a[i * m + j];
    if (sqrt(2.0 * off) <= tol) break;
    for (int p = 0; p < m; ++p)
      for (int r = p + 1; r < m; ++r) {
        const double apr = a[p * m + r];
        if (apr == 0.0) continue;
        const double tau = (a[r * m + r] - a[p * m + p]) / (2.0 * apr);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = a[k * m + p], akr = a[k * m + r];
          a[k * m + p] = c * akp - s * akr;
          a[k * m + r] = s * akp + c * akr;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a[p * m + k], ark = a[r * m + k];
          a[p * m + k] = c * apk - s * ark;
          a[r * m + k] = s * apk + c * ark;
        }
        for (int k = 0; k < m; ++k) {
          const double qkp = q[k * m + p], qkr = q[k * m + r];
          q[k * m + p] = c * qkp - s * qkr;
          q[k * m + r] = s * qkp + c * qkr;
        }
      }
  }
  for (int i = 0; i < m; ++i) ev[i] = a[i * m + i];
}

// Log of a pose (math/twist.cpp:44-67): the axis is the eigenvector of (R - I)^T (R - I) of its smallest eigenvalue
// (the right singular vector of R - I), the angle atan2(rvec . axis, trace(R) - 1).
__host__ __device__ inline void twist_log(const double* R, const double* t, double* tw) {
  double a[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[3 * i + j] = R[3 * i + j] - (i == j ? 1.0 : 0.0);
  double ata[9], ev[3], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ata[3 * i + j] = a[i] * a[j] + a[3 + i] * a[3 + j] + a[6 + i] * a[6 + j];
  small_jacobi(3, ata, ev, V);
  int k = 0;
  for (int c = 1; c < 3; ++c)
    if (ev[c] < ev[k]) k = c;
  double w[3] = {V[k], V[3 + k], V[6 + k]};
  const double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int i = 0; i < 3; ++i) w[i] /= nw;
  const double rv[3] = {a[7] - a[5], a[2] - a[6], a[3] - a[1]};
  const double wmag = atan2(rv[0] * w[0] + rv[1] * w[1] + rv[2] * w[2], a[0] + a[4] + a[8] + 2.0);
  const double h = 0.5 * wmag;
  const double wxt[3] = {w[1] * t[2] - w[2] * t[1], w[2] * t[0] - w[0] * t[2], w[0] * t[1] - w[1] * t[0]};
  const double wxwxt[3] = {w[1] * wxt[2] - w[2] * wxt[1], w[2] * wxt[0] - w[0] * wxt[2],
                           w[0] * wxt[1] - w[1] * wxt[0]};
  const double coef = fabs(wmag) < kFltEps ? 0.0 : 1.0 - h / tan(h);
  for (int i = 0; i < 3; ++i) {
    tw[i] = wmag * w[i];
    tw[3 + i] = t[i] + coef * wxwxt[i] - h * wxt[i];
  }
}

// Log(Exp(rhs) Exp(lhs)) (math/twist.h:67-73).
__host__ __device__ inline void twist_add(const double* lhs, const double* rhs, double* out) {
  double Ra[9], ta[3], Rb[9], tb[3], R[9], t[3];
  twist_exp(lhs, Ra, ta);
  twist_exp(rhs, Rb, tb);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) R[3 * i + j] = Rb[3 * i] * Ra[j] + Rb[3 * i + 1] * Ra[3 + j] + Rb[3 * i + 2] * Ra[6 + j];
    t[i] = Rb[3 * i] * ta[0] + Rb[3 * i + 1] * ta[1] + Rb[3 * i + 2] * ta[2] + tb[i];
  }
  twist_log(R, t, out);
}

// U V^T of the SVD of M with the reflection fix (common/rotation_utils.h:36-50), from the eigen-decomposition of
// M^T M with the singular values in decreasing order.
__host__ __device__ inline void nearest_rotation(const double* M, double* R) {
  double MtM[9], ev[3], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) MtM[3 * i + j] = M[i] * M[j] + M[3 + i] * M[3 + j] + M[6 + i] * M[6 + j];
  small_jacobi(3, MtM, ev, V);
  int o[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)  // stable descending order, like std::sort on a strict comparison of distinct values
    for (int j = 0; j + 1 < 3 - i; ++j)
      if (ev[o[j + 1]] > ev[o[j]]) {
        const int tmp = o[j];
        o[j] = o[j + 1];
        o[j + 1] = tmp;
      }
  double U[9], Vs[9];
  for (int k = 0; k < 3; ++k) {
    const int c = o[k];
    const double s = sqrt(ev[c] > 0.0 ? ev[c] : 0.0);
    for (int i = 0; i < 3; ++i) Vs[3 * i + k] = V[3 * i + c];
    for (int i = 0; i < 3; ++i)
      U[3 * i + k] = s > 0.0 ? (M[3 * i] * Vs[k] + M[3 * i + 1] * Vs[3 + k] + M[3 * i + 2] * Vs[6 + k]) / s : 0.0;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[3 * i + j] = U[3 * i] * Vs[3 * j] + U[3 * i + 1] * Vs[3 * j + 1] + U[3 * i + 2] * Vs[3 * j + 2];
  const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                     R[2] * (R[3] * R[7] - R[4] * R[6]);
  if (det < 0.0)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[3 * i + j] -= 2.0 * U[3 * i + 2] * Vs[3 * j + 2];
}

// Eigen's inverse with a determinant check, 2x2 and 3x3 (row-major): false (nothing written) when |det| <= thr.
__host__ __device__ inline bool inverse2(const double* m, double* out, double thr) {
  const double det = m[0] * m[3] - m[1] * m[2];
  if (!(fabs(det) > thr)) return false;
  out[0] = m[3] / det;
  out[1] = -m[1] / det;
  out[2] = -m[2] / det;
  out[3] = m[0] / det;
  return true;
}
__host__ __device__ inline bool inverse3(const double* m, double* out, double thr) {
  const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
  const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  if (!(fabs(det) > thr)) return false;
  const double c10 = m[2] * m[7] - m[1] * m[8], c11 = m[0] * m[8] - m[2] * m[6], c12 = m[1] * m[6] - m[0] * m[7];
  const double c20 = m[1] * m[5] - m[2] * m[4], c21 = m[2] * m[3] - m[0] * m[5], c22 = m[0] * m[4] - m[1] * m[3];
  const double cof[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = cof[3 * j + i] / det;
  return true;
}

// The control state of one window's solve, shared by the workgroup (LDS on the GPU).
struct MonoShared {
  double residual, old, lambda, red[kMonoThreads];
  int32_t count, new_values, failed, feasible, iteration, ok, total, n, stop, chol, ired[kMonoThreads];
};

// ---- the solver: one workgroup per window -------------------------------------------------------------------------

struct MonoSolve {
  const MonoWin& w;
  MonoShared& sh;
  int tid, nth;
  double* lds = nullptr;  // workgroup memory for the eigen-decomposition (GPU), lds_cap doubles
  int64_t lds_cap = 0;

  // workgroup sum of v in a fixed order: on the GPU lane l of the first wavefront adds the partials of threads
  // l, l + 64, l + 128, l + 192 and the 64 lane sums are added by a fixed butterfly; on the host (one thread) the
  // partial itself
  __host__ __device__ double sum(double v) {
    sh.red[tid] = v;
    MONO_SYNC();
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(kMonoThreads == 256, "the reduction below covers 256 threads");
    if (tid < 64) {
      double s = (sh.red[tid] + sh.red[tid + 64]) + (sh.red[tid + 128] + sh.red[tid + 192]);
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (tid == 0) sh.red[0] = s;
    }
#else
    if (tid == 0) {
      double s = 0.0;
      for (int i = 0; i < nth; ++i) s += sh.red[i];
      sh.red[0] = s;
    }
#endif
    MONO_SYNC();
    const double s = sh.red[0];
    MONO_SYNC();
    return s;
  }
  __host__ __device__ int isum(int v) {
    sh.ired[tid] = v;
    MONO_SYNC();
    if (tid == 0) {
      int s = 0;
      for (int i = 0; i < nth; ++i) s += sh.ired[i];
      sh.ired[0] = s;
    }
    MONO_SYNC();
    const int s = sh.ired[0];
    MONO_SYNC();
    return s;
  }

  // Sliced sums: `items` independent sums (frames, reduced-system blocks), each over a list of terms, are split
  // into `slices` strided slices per item (slice s takes terms s, s + slices, ...); thread (item, slice) adds its
  // terms into kPart partials, and the partials of an item are then added in slice order. The result depends on the
  // workgroup size only through `slices`, which is fixed for a launch.
  __host__ __device__ int slices_for(int items) const {
    const int s = items > 0 ? nth / items : 1;
    return s < 1 ? 1 : s;
  }

  // frames' poses from their twists (Exp(a)) and, with `nearest`, the nearest rotations
  __host__ __device__ void frame_poses(bool nearest) {
    for (int f = tid; f < w.nF; f += nth) {
      twist_exp(w.fa + 6 * f, w.fM + 12 * f, w.fM + 12 * f + 9);
      if (nearest) nearest_rotation(w.fM + 12 * f, w.fRn + 9 * f);
    }
    MONO_SYNC();
  }

  __host__ __device__ void apply(int f, const double* b, double* q) const {
    const double* M = w.fM + 12 * f;
    for (int i = 0; i < 3; ++i) q[i] = M[3 * i] * b[0] + M[3 * i + 1] * b[1] + M[3 * i + 2] * b[2] + M[9 + i];
  }

  __host__ __device__ void filter_hither() {  // mono_sba_solver.cpp:346-361
    frame_poses(false);
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t]) continue;
      for (int q = w.track_obs[t]; q < w.track_obs[t + 1]; ++q) {
        double p[3];
        apply(w.oframe[w.tobs[q]], w.tb + 3 * t, p);
        if (!(p[2] > kHither)) {
          w.tskip[t] = 1;
          break;
        }
      }
    }
    MONO_SYNC();
  }

  __host__ __device__ void update_aDbD() {  // mono_sba_solver.cpp:770-786
    for (int f = tid; f < w.nF; f += nth)
      if (!w.constrained[f]) twist_add(w.fa + 6 * f, w.fda + 6 * f, w.faD + 6 * f);
    for (int t = tid; t < w.nT; t += nth)
      if (!w.tskip[t])
        for (int d = 0; d < 3; ++d) w.tbD[3 * t + d] = w.tb[3 * t + d] + w.tdb[3 * t + d];
    MONO_SYNC();
  }

  __host__ __device__ void update() {  // mono_sba_solver.cpp:788-810
    for (int f = tid; f < w.nF; f += nth)
      if (!w.constrained[f])
        for (int d = 0; d < 6; ++d) w.fa[6 * f + d] = w.faD[6 * f + d];
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t])
        w.tskip[t] = 0;
      else
        for (int d = 0; d < 3; ++d) w.tb[3 * t + d] = w.tbD[3 * t + d];
    }
    MONO_SYNC();
  }

  // the mean Huber term over the observations of the tracks not skipped, from aD and bD (mono_sba_solver.cpp:743-768);
  // FLT_MAX when there is none
  __host__ __device__ double calc_residual() {
    for (int f = tid; f < w.nF; f += nth) twist_exp(w.faD + 6 * f, w.fM + 12 * f, w.fM + 12 * f + 9);
    MONO_SYNC();
    double r = 0.0;
    int total = 0;
    for (int o = tid; o < w.nO; o += nth) {
      const int t = w.otrack[o];
      if (w.tskip[t]) continue;
      double p[3];
      apply(w.oframe[o], w.tbD + 3 * t, p);
      const double z = avoid_zero(p[2]);
      w.oerr[2 * o] = double(w.oxy[2 * o]) - p[0] / z;
      w.oerr[2 * o + 1] = double(w.oxy[2 * o + 1]) - p[1] / z;
      r += huber(quad(w.oerr + 2 * o, w.tinfo + 4 * t));
      ++total;
    }
    const double rs = sum(r);
    const int n = isum(total);
    return n != 0 ? rs / double(n) : kFltMax;
  }

  // per-track information from the covariance of its residuals over its frames (mono_sba_solver.cpp:363-432):
  // variances floored at 1e-5; with more than three samples the cross term, then an LDLT with the floor on D,
  // reassembled without the pivot permutation; inverted when the determinant exceeds FLT_EPSILON, else the
  // reciprocal diagonal
  __host__ __device__ void calc_information() {
    const double thresh = double(1e-5f);
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t]) continue;
      double mx = 0, vx = 0, my = 0, vy = 0, mxy = 0, vxy = 0;
      int cnt = 0;
      for (int q = w.track_obs[t]; q < w.track_obs[t + 1]; ++q) {
        const int o = w.tobs[q];
        const double ex = w.oerr[2 * o], ey = w.oerr[2 * o + 1], exy = ex * ey;
        ++cnt;
        const double dx = ex - mx, dy = ey - my, dxy = exy - mxy;
        mx += dx / double(cnt);
        my += dy / double(cnt);
        mxy += dxy / double(cnt);
        vx += dx * dx * double(cnt - 1) / double(cnt);
        vy += dy * dy * double(cnt - 1) / double(cnt);
        vxy += dxy * dxy * double(cnt - 1) / double(cnt);
      }
      (void)vxy;
      const double varx = vx / double(cnt), vary = vy / double(cnt);
      double c[4] = {varx < thresh ? thresh : varx, 0.0, 0.0, vary < thresh ? thresh : vary};
      if (cnt > 3) {
        c[1] = c[2] = mxy - mx * my;
        const bool swap = fabs(c[3]) > fabs(c[0]);
        const double p00 = swap ? c[3] : c[0], p11 = swap ? c[0] : c[3], p01 = c[1];
        const double d0 = p00, l = p01 / p00, d1 = p11 - l * p01;
        const double e0 = d0 > thresh ? d0 : thresh, e1 = d1 > thresh ? d1 : thresh;
        c[0] = e0;
        c[1] = c[2] = l * e0;
        c[3] = l * l * e0 + e1;
      }
      double* L = w.tinfo + 4 * t;
      if (!inverse2(c, L, kFltEps)) {
        L[0] = 1.0 / c[0];
        L[3] = 1.0 / c[3];
        L[1] = L[2] = 0.0;
      }
    }
    MONO_SYNC();
  }

  __host__ __device__ void compute_weights() {  // mono_sba_solver.cpp:435-444 (every track, skipped or not)
    for (int o = tid; o < w.nO; o += nth) w.ow[o] = dhuber(quad(w.oerr + 2 * o, w.tinfo + 4 * w.otrack[o]));
    MONO_SYNC();
  }

  // projection derivatives of every observation of a track not skipped (mono_sba_solver.cpp:704-741,
  // math/twist.h:152-206): B with the nearest rotation of the frame, A = -(the reference's camera derivatives)
  __host__ __device__ void calc_derivatives() {
    frame_poses(true);
    for (int o = tid; o < w.nO; o += nth) {
      const int t = w.otrack[o], f = w.oframe[o];
      if (w.tskip[t]) continue;
      double q[3];
      apply(f, w.tb + 3 * t, q);
      const double iz = 1.0 / avoid_zero(q[2]);
      const double x = q[0] * iz, y = q[1] * iz;
      const double* Rn = w.fRn + 9 * f;
      double* B = w.oB + 6 * o;
      for (int j = 0; j < 3; ++j) {
        B[j] = (Rn[j] - Rn[6 + j] * x) * iz;
        B[3 + j] = (Rn[3 + j] - Rn[6 + j] * y) * iz;
      }
      if (!w.constrained[f]) {
        const double am[12] = {x * y, -x * x - 1.0, y, -iz, 0.0, x * iz, y * y + 1.0, -(x * y), -x, 0.0, -iz, y * iz};
        for (int e = 0; e < 12; ++e) w.oA[12 * o + e] = -am[e];
      }
    }
    MONO_SYNC();
  }

  // U and errA per free frame (its observations in track order), V and errB per track, W per observation
  // (mono_sba_solver.cpp:586-678)
  __host__ __device__ void calc_UVW() {
    // U and errA of every frame, sliced over its observations
    const int fs = slices_for(w.nF);
    for (int it = tid; it < w.nF * fs; it += nth) {
      const int f = it / fs, sl = it % fs;
      double P[kPart] = {};
      if (!w.constrained[f]) {
        for (int q = w.frame_obs[f] + sl; q < w.frame_obs[f + 1]; q += fs) {
          const int o = w.fobs[q], t = w.otrack[o];
          if (w.tskip[t]) continue;
          const double* A = w.oA + 12 * o;
          const double* L = w.tinfo + 4 * t;
          const double wo = w.ow[o];
          double tmp[6][2];
          for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 2; ++c) tmp[r][c] = A[r] * (wo * L[c]) + A[6 + r] * (wo * L[2 + c]);
          for (int r = 0; r < 6; ++r) {
            for (int c = 0; c < 6; ++c) P[6 * r + c] += tmp[r][0] * A[c] + tmp[r][1] * A[6 + c];
            P[36 + r] += tmp[r][0] * w.oerr[2 * o] + tmp[r][1] * w.oerr[2 * o + 1];
          }
        }
      }
      for (int i = 0; i < kPart; ++i) w.part[kPart * it + i] = P[i];
    }
    MONO_SYNC();
    for (int it = tid; it < w.nF * kPart; it += nth) {
      const int f = it / kPart, i = it % kPart;
      if (w.constrained[f]) continue;
      double v = 0.0;
      for (int sl = 0; sl < fs; ++sl) v += w.part[kPart * (f * fs + sl) + i];
      if (i < 36)
        w.fUo[36 * f + i] = v;
      else
        w.ferrA[6 * f + i - 36] = v;
    }
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t]) continue;
      double V[9] = {}, e[3] = {};
      const double* L = w.tinfo + 4 * t;
      for (int q = w.track_obs[t]; q < w.track_obs[t + 1]; ++q) {
        const int o = w.tobs[q];
        const double* B = w.oB + 6 * o;
        const double wo = w.ow[o];
        double tmp[3][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 2; ++c) tmp[r][c] = B[r] * (wo * L[c]) + B[3 + r] * (wo * L[2 + c]);
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) V[3 * r + c] += tmp[r][0] * B[c] + tmp[r][1] * B[3 + c];
          e[r] += tmp[r][0] * w.oerr[2 * o] + tmp[r][1] * w.oerr[2 * o + 1];
        }
      }
      for (int i = 0; i < 9; ++i) w.tV[9 * t + i] = V[i];
      for (int i = 0; i < 3; ++i) w.terrB[3 * t + i] = e[i];
    }
    for (int o = tid; o < w.nO; o += nth) {
      const int t = w.otrack[o];
      if (w.tskip[t] || w.constrained[w.oframe[o]]) continue;
      const double* A = w.oA + 12 * o;
      const double* B = w.oB + 6 * o;
      const double* L = w.tinfo + 4 * t;
      const double wo = w.ow[o];
      const double Li[4] = {L[0] * wo, L[1] * wo, L[2] * wo, L[3] * wo};
      for (int r = 0; r < 6; ++r) {
        const double m0 = A[r] * Li[0] + A[6 + r] * Li[2];
        const double m1 = A[r] * Li[1] + A[6 + r] * Li[3];
        for (int c = 0; c < 3; ++c) w.oW[18 * o + 3 * r + c] = m0 * B[c] + m1 * B[3 + c];
      }
    }
    MONO_SYNC();
  }

  // damped U and V, V's inverse with the determinant check (a failed inverse skips the track), Y = W V*^-1
  // (mono_sba_solver.cpp:446-510, 892-899)
  __host__ __device__ void step_UV() {
    const double lam = sh.lambda;
    for (int f = tid; f < w.nF; f += nth) {
      if (w.constrained[f]) continue;
      for (int i = 0; i < 36; ++i) w.fUa[36 * f + i] = w.fUo[36 * f + i];
      for (int i = 0; i < 6; ++i) w.fUa[36 * f + 7 * i] *= 1.0 + lam;
    }
    const double l3 = lam * lam * lam;
    const double thr = 1.0 / (double(1e6f) > l3 ? double(1e6f) : l3);
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t]) continue;
      double V[9];
      for (int i = 0; i < 9; ++i) V[i] = w.tV[9 * t + i];
      for (int i = 0; i < 3; ++i) V[4 * i] *= 1.0 + lam;
      if (!inverse3(V, w.tinvVa + 9 * t, thr)) w.tskip[t] = 1;
    }
    MONO_SYNC();
    for (int o = tid; o < w.nO; o += nth) {
      const int t = w.otrack[o];
      if (w.tskip[t] || w.constrained[w.oframe[o]]) continue;
      const double* Wo = w.oW + 18 * o;
      const double* X = w.tinvVa + 9 * t;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c)
          w.oY[18 * o + 3 * r + c] = Wo[3 * r] * X[c] + Wo[3 * r + 1] * X[3 + c] + Wo[3 * r + 2] * X[6 + c];
    }
    MONO_SYNC();
  }

  // Parallel cyclic Jacobi (Brent-Luk round-robin ordering: each step rotates n/2 disjoint index pairs at once) of
  // the n x n symmetric E into eigenvalues ev and eigenvectors Q; sweeps until the off-diagonal norm falls to
  // eps ||E||_F (Golub & Van Loan, section 8.5). A step computes its n/2 rotations from its input E, then the
  // thread of the 2 x 2 block (pair k rows, pair k2 columns) writes J_k^T E J_k2 of its block to the other buffer
  // and the thread of (row r, pair k) rotates Q's row r in place. E and Q live in workgroup memory when 3 n^2
  // doubles fit in it, else in the window's work area.
  __host__ __device__ void rotation(const double* E, int n, int k, int step, int m, int& a, int& b, double& c,
                                    double& s) const {
    a = k == 0 ? 0 : 1 + (k - 1 + step) % (m - 1);
    b = 1 + (m - 2 - k + step) % (m - 1);
    if (a > b) {
      const int tmp = a;
      a = b;
      b = tmp;
    }
    c = 1.0;
    s = 0.0;
    if (b < n) {
      const double apr = E[a * n + b];
      if (apr != 0.0) {
        const double tau = (E[b * n + b] - E[a * n + a]) / (2.0 * apr);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        c = 1.0 / sqrt(1.0 + t * t);
        s = t * c;
      }
    }
  }
  __host__ __device__ void eigen(int n) {
    const int nn = n * n;
    const bool in_lds = lds != nullptr && 3 * static_cast<int64_t>(nn) <= lds_cap;
    double* E0 = in_lds ? lds : w.E;
    double* E1 = in_lds ? lds + nn : w.E2;
    double* Q = in_lds ? lds + 2 * nn : w.Q;
    double* rot = in_lds && 3 * static_cast<int64_t>(nn) + 2 * (n + 2) <= lds_cap ? lds + 3 * nn : w.rot;
    for (int i = tid; i < nn; i += nth) {
      if (in_lds) E0[i] = w.E[i];
      Q[i] = (i / n == i % n) ? 1.0 : 0.0;
    }
    MONO_SYNC();
    double fr = 0.0;
    for (int i = tid; i < nn; i += nth) fr += E0[i] * E0[i];
    const double fro = sum(fr);
    const double tol = 2.220446049250313e-16 * sqrt(fro);
    const int m = n + (n & 1);  // even number of slots; slot m-1 is a dummy when n is odd
    const int h = m / 2;
    for (int sweep = 0; sweep < 60; ++sweep) {
      double off = 0.0;
      for (int i = tid; i < nn; i += nth)
        if (i % n > i / n) off += E0[i] * E0[i];
      if (sqrt(2.0 * sum(off)) <= tol) break;
      for (int step = 0; step < m - 1; ++step) {
        // the step's rotations, one per pair (slot indexes and cosine, sine), from the step's input E
        for (int k = tid; k < h; k += nth) {
          int a, b;
          double c, s;
          rotation(E0, n, k, step, m, a, b, c, s);
          rot[4 * k] = a;
          rot[4 * k + 1] = b;
          rot[4 * k + 2] = c;
          rot[4 * k + 3] = s;
        }
        MONO_SYNC();
        for (int e = tid; e < h * h + n * h; e += nth) {
          int a, b;
          double c, s;
          auto rotk = [&](int k, int& ra, int& rb, double& rc, double& rs) {
            ra = int(rot[4 * k]);
            rb = int(rot[4 * k + 1]);
            rc = rot[4 * k + 2];
            rs = rot[4 * k + 3];
          };
          if (e < h * h) {
            const int k = e / h, k2 = e % h;
            int a2, b2;
            double c2, s2;
            rotk(k, a, b, c, s);
            rotk(k2, a2, b2, c2, s2);
            const bool rb = b < n, cb = b2 < n;  // the dummy slot has no row or column
            const double xaa = E0[a * n + a2], xab = cb ? E0[a * n + b2] : 0.0;
            const double xba = rb ? E0[b * n + a2] : 0.0, xbb = rb && cb ? E0[b * n + b2] : 0.0;
            // columns first (E J_k2), then rows (J_k^T (E J_k2))
            const double ya_a = c2 * xaa - s2 * xab, ya_b = s2 * xaa + c2 * xab;
            const double yb_a = c2 * xba - s2 * xbb, yb_b = s2 * xba + c2 * xbb;
            E1[a * n + a2] = c * ya_a - s * yb_a;
            if (cb) E1[a * n + b2] = c * ya_b - s * yb_b;
            if (rb) E1[b * n + a2] = s * ya_a + c * yb_a;
            if (rb && cb) E1[b * n + b2] = s * ya_b + c * yb_b;
          } else {
            const int r = (e - h * h) / h, k = (e - h * h) % h;
            rotk(k, a, b, c, s);
            if (b >= n) continue;
            double* Qr = Q + r * n;
            const double qa = Qr[a], qb = Qr[b];
            Qr[a] = c * qa - s * qb;
            Qr[b] = s * qa + c * qb;
          }
        }
        MONO_SYNC();
        double* t = E0;
        E0 = E1;
        E1 = t;
      }
    }
    for (int i = tid; i < n; i += nth) w.ev[i] = E0[i * n + i];
    if (in_lds)
      for (int i = tid; i < nn; i += nth) w.Q[i] = Q[i];
    MONO_SYNC();
  }

  // In-place Cholesky A = L L^T (lower triangle) of the n x n row-major A, right-looking by columns; false on a
  // non-positive pivot.
  __host__ __device__ bool cholesky(int n, double* A) {
    if (tid == 0) sh.chol = 1;
    MONO_SYNC();
    for (int k = 0; k < n; ++k) {
      if (tid == 0) {
        const double d = A[k * n + k];
        if (!(d > 0.0)) sh.chol = 0;
        A[k * n + k] = d > 0.0 ? sqrt(d) : 1.0;
      }
      MONO_SYNC();
      if (!sh.chol) return false;
      const double dk = A[k * n + k];
      for (int i = k + 1 + tid; i < n; i += nth) A[i * n + k] /= dk;
      MONO_SYNC();
      const int m = n - k - 1;  // trailing lower triangle, row-major over (i, j), k < j <= i
      for (int e = tid; e < m * (m + 1) / 2; e += nth) {
        int i = static_cast<int>((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        const int j = e - i * (i + 1) / 2;
        A[(k + 1 + i) * n + (k + 1 + j)] -= A[(k + 1 + i) * n + k] * A[(k + 1 + j) * n + k];
      }
      MONO_SYNC();
    }
    return true;
  }

  // The reference solves the reduced system through a rank-revealing decomposition with the rank threshold
  // n FLT_EPSILON max |eigenvalue|. Full rank is certified without it when S - tau I has a Cholesky factor, with
  // tau = n FLT_EPSILON times the largest absolute row sum of S (an upper bound of max |eigenvalue|): then every
  // eigenvalue exceeds tau >= the threshold, the pseudo-inverse is the inverse, and S is factored by Cholesky.
  // Returns the factor of S (in workgroup memory when it fits), or null when the certificate fails (the caller then
  // takes the eigen-decomposition).
  __host__ __device__ double* certified_factor(int n) {
    const int nn = n * n;
    double* A = lds != nullptr && static_cast<int64_t>(nn) <= lds_cap ? lds : w.E2;
    double rmax = 0.0;
    for (int i = tid; i < n; i += nth) {
      double r = 0.0;
      for (int j = 0; j < n; ++j) r += fabs(w.S[i * n + j]);
      rmax = r > rmax ? r : rmax;
    }
    sh.red[tid] = rmax;
    MONO_SYNC();
    if (tid == 0) {
      double m = 0.0;
      for (int i = 0; i < nth; ++i) m = sh.red[i] > m ? sh.red[i] : m;
      sh.red[0] = m;
    }
    MONO_SYNC();
    const double tau = double(n) * kFltEps * sh.red[0];
    MONO_SYNC();
    for (int e = tid; e < nn; e += nth) A[e] = w.S[e] - (e / n == e % n ? tau : 0.0);
    MONO_SYNC();
    if (!cholesky(n, A)) return nullptr;
    for (int e = tid; e < nn; e += nth) A[e] = w.S[e];
    MONO_SYNC();
    if (!cholesky(n, A)) return nullptr;
    return A;
  }

  // out = (L L^T)^-1 rhs, forward and backward substitution by one thread
  __host__ __device__ void chol_solve(int n, const double* L, const double* rhs, double* out) {
    if (tid == 0) {
      for (int i = 0; i < n; ++i) {
        double v = rhs[i];
        for (int j = 0; j < i; ++j) v -= L[i * n + j] * out[j];
        out[i] = v / L[i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double v = out[i];
        for (int j = i + 1; j < n; ++j) v -= L[j * n + i] * out[j];
        out[i] = v / L[i * n + i];
      }
    }
    MONO_SYNC();
  }

  // out = Q diag(1/ev) Q^T rhs
  __host__ __device__ void psolve(int n, const double* rhs, double* out) {
    for (int k = tid; k < n; k += nth) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += w.Q[i * n + k] * rhs[i];
      w.rot[k] = s / w.ev[k];
    }
    MONO_SYNC();
    for (int i = tid; i < n; i += nth) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += w.Q[i * n + k] * w.rot[k];
      out[i] = s;
    }
    MONO_SYNC();
  }

  __host__ __device__ double norm(int n, const double* v) {
    double s = 0.0;
    for (int i = tid; i < n; i += nth) s += v[i] * v[i];
    return sqrt(sum(s));
  }

  // The reduced system, its solve and the back-substitution (mono_sba_solver.cpp:681-702, 812-890): false when no
  // frame is free or the system is rank deficient.
  __host__ __device__ bool calc_adjustments() {
    const double lam = sh.lambda;
    const int nf = w.nFree, n = 6 * nf;
    // eps = errA - sum over the frame's observations of Y errB, sliced over the observations
    const int fs = slices_for(w.nF);
    for (int it = tid; it < w.nF * fs; it += nth) {
      const int f = it / fs, sl = it % fs;
      double P[6] = {};
      if (!w.constrained[f]) {
        for (int q = w.frame_obs[f] + sl; q < w.frame_obs[f + 1]; q += fs) {
          const int o = w.fobs[q], t = w.otrack[o];
          if (w.tskip[t]) continue;
          const double* Y = w.oY + 18 * o;
          const double* eb = w.terrB + 3 * t;
          for (int r = 0; r < 6; ++r) P[r] += Y[3 * r] * eb[0] + Y[3 * r + 1] * eb[1] + Y[3 * r + 2] * eb[2];
        }
      }
      for (int r = 0; r < 6; ++r) w.part[kPart * it + r] = P[r];
    }
    MONO_SYNC();
    for (int it = tid; it < 6 * w.nF; it += nth) {
      const int f = it / 6, r = it % 6;
      if (w.constrained[f]) continue;
      double v = 0.0;
      for (int sl = 0; sl < fs; ++sl) v += w.part[kPart * (f * fs + sl) + r];
      w.feps[6 * f + r] = w.ferrA[6 * f + r] - v;
    }
    MONO_SYNC();
    if (n == 0) {
      for (int t = tid; t < w.nT; t += nth)
        for (int d = 0; d < 3; ++d) w.tdb[3 * t + d] = 0.0;
      MONO_SYNC();
      return false;
    }
    // S = blockdiag(Ua) - sum over tracks of Y_j W_k^T for every ordered pair of its free frames, each block's sum
    // sliced over its track pairs; b = eps
    const int bs = slices_for(nf * nf);
    for (int it = tid; it < nf * nf * bs; it += nth) {
      const int bi = it / bs, sl = it % bs;
      double P[36] = {};
      for (int q = w.blk[bi] + sl; q < w.blk[bi + 1]; q += bs) {
        const int oj = w.bpair[q].x, ok = w.bpair[q].y;
        if (w.tskip[w.otrack[oj]]) continue;
        double Y[18], W[18];
        for (int i = 0; i < 18; ++i) {
          Y[i] = w.oY[18 * oj + i];
          W[i] = w.oW[18 * ok + i];
        }
        for (int rr = 0; rr < 6; ++rr)
          for (int cc = 0; cc < 6; ++cc)
            P[6 * rr + cc] += Y[3 * rr] * W[3 * cc] + Y[3 * rr + 1] * W[3 * cc + 1] + Y[3 * rr + 2] * W[3 * cc + 2];
      }
      for (int i = 0; i < 36; ++i) w.part[kPart * it + i] = P[i];
    }
    MONO_SYNC();
    for (int e = tid; e < n * n; e += nth) {
      const int r = e / n, c = e % n, j = r / 6, k = c / 6, rr = r % 6, cc = c % 6;
      double v = 0.0;
      if (j == k) {
        for (int f = 0; f < w.nF; ++f)
          if (w.free_idx[f] == j) v = w.fUa[36 * f + 6 * rr + cc];
      }
      const int bi = j * nf + k;
      double p = 0.0;
      for (int sl = 0; sl < bs; ++sl) p += w.part[kPart * (bi * bs + sl) + 6 * rr + cc];
      v -= p;
      w.S[e] = v;
      w.E[e] = v;
    }
    for (int f = tid; f < w.nF; f += nth)
      if (!w.constrained[f])
        for (int r = 0; r < 6; ++r) w.Sb[6 * w.free_idx[f] + r] = w.feps[6 * f + r];
    MONO_SYNC();
    // the last free frame's block with its diagonal divided by (1 + lambda), for the covariance
    if (tid == 0) {
      const int jj = n - 6;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) w.prep[r * 6 + c] = w.S[(jj + r) * n + jj + c];
      for (int r = 0; r < 6; ++r) w.prep[r * 7] /= 1.0 + lam;
    }
    MONO_SYNC();
    double* L = certified_factor(n);
    if (!L) {
      eigen(n);
      // rank: |ev_k| > n FLT_EPSILON max |ev| for every k
      if (tid == 0) {
        double smax = 0.0;
        for (int k = 0; k < n; ++k) smax = fabs(w.ev[k]) > smax ? fabs(w.ev[k]) : smax;
        const double thr = double(n) * kFltEps * smax;
        int rank = 0;
        for (int k = 0; k < n; ++k) rank += fabs(w.ev[k]) > thr;
        sh.ok = rank == n;
      }
      MONO_SYNC();
      if (!sh.ok) return false;
    }
    // IterativeSolver: up to five steps x += solve(b - S x), kept while the residual norm decreases
    for (int i = tid; i < n; i += nth) {
      w.xt[i] = 0.0;
      w.bd[i] = w.Sb[i];
      w.x[i] = 0.0;
    }
    MONO_SYNC();
    double min_norm = 1.79769313486231570e+308;
    double nrm = norm(n, w.Sb);
    for (int it = 0; nrm < min_norm && it < 5; ++it) {
      min_norm = nrm;
      for (int i = tid; i < n; i += nth) w.x[i] = w.xt[i];
      MONO_SYNC();
      if (L)
        chol_solve(n, L, w.bd, w.dv);
      else
        psolve(n, w.bd, w.dv);
      for (int i = tid; i < n; i += nth) w.xt[i] += w.dv[i];
      MONO_SYNC();
      for (int i = tid; i < n; i += nth) {
        double s = w.Sb[i];
        for (int j = 0; j < n; ++j) s -= w.S[i * n + j] * w.xt[j];
        w.bd[i] = s;
      }
      MONO_SYNC();
      nrm = norm(n, w.bd);
      if (nrm < kFltEps) {
        for (int i = tid; i < n; i += nth) w.x[i] = w.xt[i];
        MONO_SYNC();
        break;
      }
    }
    for (int f = tid; f < w.nF; f += nth)
      for (int i = 0; i < 6; ++i) w.fda[6 * f + i] = w.constrained[f] ? 0.0 : w.x[6 * w.free_idx[f] + i];
    MONO_SYNC();
    for (int t = tid; t < w.nT; t += nth) {
      if (w.tskip[t]) continue;
      double wda[3] = {0.0, 0.0, 0.0};
      for (int q = w.track_obs[t]; q < w.track_obs[t + 1]; ++q) {
        const int o = w.tobs[q], f = w.oframe[o];
        if (w.constrained[f]) continue;
        const double* W = w.oW + 18 * o;
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 6; ++r) wda[c] += W[3 * r + c] * w.fda[6 * f + r];
      }
      const double* eb = w.terrB + 3 * t;
      const double pt[3] = {eb[0] - wda[0], eb[1] - wda[1], eb[2] - wda[2]};
      const double* X = w.tinvVa + 9 * t;
      for (int r = 0; r < 3; ++r) w.tdb[3 * t + r] = X[3 * r] * pt[0] + X[3 * r + 1] * pt[1] + X[3 * r + 2] * pt[2];
    }
    MONO_SYNC();
    return true;
  }

  // CompleteOrthogonalDecomposition(prep).pseudoInverse() of the symmetric part of the kept block, one thread
  __host__ __device__ void covariance_from_prep() {
    if (tid == 0) {
      double P[36], ev[6], q[36];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) P[6 * r + c] = 0.5 * (w.prep[6 * r + c] + w.prep[6 * c + r]);
      small_jacobi(6, P, ev, q);
      double smax = 0.0;
      for (int k = 0; k < 6; ++k) smax = fabs(ev[k]) > smax ? fabs(ev[k]) : smax;
      for (int i = 0; i < 36; ++i) w.cov[i] = 0.0;
      if (smax > 0.0)
        for (int k = 0; k < 6; ++k) {
          if (!(fabs(ev[k]) > 6.0 * kFltEps * smax)) continue;
          const double inv = 1.0 / ev[k];
          for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) w.cov[6 * i + j] += q[6 * i + k] * inv * q[6 * j + k];
        }
    }
    MONO_SYNC();
  }

  // MonoSBASolver::solve (mono_sba_solver.cpp:205-301)
  __host__ __device__ void run() {
    // start: twists from the input poses (a maps world to frame coordinates), points, identity information
    for (int f = tid; f < w.nF; f += nth) {
      const float* T = w.pose_in + 16 * f;  // world_from_camera, column-major
      double R[9], t[3];
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) R[3 * i + j] = T[4 * i + j];  // transpose of the rotation
        t[i] = -(double(T[4 * i]) * T[12] + double(T[4 * i + 1]) * T[13] + double(T[4 * i + 2]) * T[14]);
      }
      twist_log(R, t, w.fa + 6 * f);
      for (int d = 0; d < 6; ++d) {
        w.fda[6 * f + d] = 0.0;
        w.ferrA[6 * f + d] = 0.0;
        w.faD[6 * f + d] = w.constrained[f] ? w.fa[6 * f + d] : 0.0;
      }
    }
    for (int t = tid; t < w.nT; t += nth) {
      for (int d = 0; d < 3; ++d) {
        w.tb[3 * t + d] = w.pts_in[3 * t + d];
        w.tdb[3 * t + d] = 0.0;
      }
      w.tskip[t] = 0;
      double* L = w.tinfo + 4 * t;
      L[0] = L[3] = 1.0;
      L[1] = L[2] = 0.0;
    }
    for (int o = tid; o < w.nO; o += nth) {
      w.oerr[2 * o] = w.oerr[2 * o + 1] = 0.0;
      w.ow[o] = 1.0;
    }
    for (int i = tid; i < 8; i += nth) w.out_scalars[i] = 0.0;
    for (int i = tid; i < 36; i += nth) w.prep[i] = w.cov[i] = 0.0;
    if (tid == 0) {
      sh.lambda = double(0.001f);
      sh.count = w.max_iterations;
      sh.feasible = 1;
      sh.failed = 0;
      sh.iteration = 0;
    }
    MONO_SYNC();
    filter_hither();
    update_aDbD();
    calc_residual();
    calc_information();
    compute_weights();
    const double r0 = calc_residual();
    if (tid == 0) sh.residual = r0;
    MONO_SYNC();
    if (r0 == kFltMax) {
      for (int i = tid; i < 36; i += nth) w.cov[i] = i % 7 == 0 ? 1.0 : 0.0;
      if (tid == 0) {
        sh.feasible = 0;
        sh.iteration = 0;
      }
      MONO_SYNC();
      finish(0.0);
      return;
    }
    if (tid == 0) {
      sh.old = r0;
      sh.new_values = 1;
    }
    MONO_SYNC();
    const double residual_initial = r0;
    while (w.min_residual <= sh.residual && sh.count >= 0 && sh.lambda <= kLambdaMax) {
      if (sh.new_values) {
        calc_derivatives();
        calc_UVW();
      }
      step_UV();
      const bool adj = calc_adjustments();
      if (adj) {
        update_aDbD();
        const double r = calc_residual();
        if (tid == 0) sh.residual = r;
        MONO_SYNC();
        if (r == kFltMax) {
          if (tid == 0) sh.feasible = 0;
          MONO_SYNC();
          finish(residual_initial);
          return;
        }
      } else {
        if (sh.residual == kFltMax) {
          if (tid == 0) sh.feasible = 0;
          MONO_SYNC();
          finish(residual_initial);
          return;
        }
        MONO_SYNC();
        if (tid == 0) sh.residual *= 2.0;
        MONO_SYNC();
      }
      const double delta = fabs(sh.residual - sh.old);
      const bool worse = sh.residual > sh.old;
      MONO_SYNC();
      if (worse) {
        if (tid == 0) {
          sh.lambda *= 10.0;
          sh.new_values = 0;
        }
        MONO_SYNC();
      } else {
        if (tid == 0) {
          if (sh.lambda > kLambdaMin) sh.lambda /= 10.0;
          sh.old = sh.residual;
          sh.new_values = 1;
        }
        MONO_SYNC();
        update();
        filter_hither();
        calc_information();
        compute_weights();
        covariance_from_prep();
      }
      const bool stop = w.min_residual_speed * sh.residual > delta;
      MONO_SYNC();
      if (stop) break;
      if (tid == 0) --sh.count;
      MONO_SYNC();
    }
    if (tid == 0) {
      sh.residual = sh.old;
      sh.failed = sh.lambda > kLambdaMax || sh.count < 0;
      sh.iteration = w.max_iterations - sh.count;
    }
    MONO_SYNC();
    if (sh.failed) covariance_from_prep();
    finish(residual_initial);
  }

  // world_from_camera of every frame (the inverse of Exp(a)), column-major, and the points
  __host__ __device__ void export_state(float* poses_out, float* pts_out) {
    for (int f = tid; f < w.nF; f += nth) {
      double R[9], t[3];
      twist_exp(w.fa + 6 * f, R, t);
      float* T = poses_out + 16 * f;
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T[4 * j + i] = float(R[3 * j + i]);  // (R^T)(i, j)
        T[12 + i] = float(-(R[i] * t[0] + R[3 + i] * t[1] + R[6 + i] * t[2]));
        T[4 * i + 3] = 0.f;
      }
      T[15] = 1.f;
    }
    for (int i = tid; i < 3 * w.nT; i += nth) pts_out[i] = float(w.tb[i]);
    MONO_SYNC();
  }

  __host__ __device__ void finish(double residual_initial) {
    if (tid == 0) {
      w.out_scalars[0] = sh.residual;
      w.out_scalars[1] = residual_initial;
      w.out_scalars[2] = sh.lambda;
      w.out_scalars[3] = sh.iteration;
      w.out_scalars[4] = sh.failed;
      w.out_scalars[5] = sh.feasible;
    }
    MONO_SYNC();
  }
};

}  // namespace sba_mono
