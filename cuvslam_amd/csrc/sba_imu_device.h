// Device-side data of one inertial window (shared by the HIP kernels in sba_imu.hip and the host runtime).
//
// Every quantity is FP64: the window's state (points, and per key frame R and t of w_from_imu, velocity and both
// biases), the observations, the preintegrations, the model (residuals and Jacobians), the full system (U dense
// over the free key frames' 15 variables, V and W blocks per point and per (point, key frame) pair), the reduced
// system and the step. Streams that the host packs:
//   pobs      observations ordered by (point, key frame, camera); point_obs[i] .. point_obs[i+1] are point i's
//   pairs     one per distinct (point, free key frame); pair_obs[2p] .. pair_obs[2p+1] its observations in pobs
//             order, pair_pose its free key frame (0-based among the free ones), point_pair[i] .. point_pair[i+1]
//             point i's pairs (by key frame)
//   kobs      observations of each free key frame in input order: pose_obs[j] .. pose_obs[j+1]
//   kpair     pairs of each free key frame in point order: pose_pair[j] .. pose_pair[j+1]
//   blocks    the lower 15x15 blocks (j >= k) of the reduced system, with the (pair of j, pair of k) products of
//             the points that both key frames observe: blk_tri[b] .. blk_tri[b+1] into tri (int2)
#pragma once

#include <stdint.h>

namespace sba_imu_dev {

constexpr int kThreads = 256;
constexpr int kEdgeThreads = 256;  // one workgroup per inertial edge
constexpr int kReducedThreads = 1024, kReducedSlices = 28;  // 36 x 28 partial sums per reduced block
constexpr int kPoseFullThreads = 1024;  // 42 values x 24 slices of a key frame's reprojection sums
constexpr int kFactorThreads = 512;
constexpr int kMaxFree = 64;  // free key frames of one window (reduced system order <= 960)
// one key frame's state: R (9, row-major), t (3), v (3), bg (3), ba (3)
constexpr int kPS = 21;
constexpr int kT = 9, kV = 12, kBg = 15, kBa = 18;

// The preintegration of key frame i's inertial edge as the kernels read it (row-major matrices).
struct Preint {
  double dR[9], JRg[9], JVg[9], JVa[9], JPg[9], JPa[9];
  double dV[3], dP[3];
  double bg0[3], ba0[3];  // biases of the integration
  double dbg[3];          // updated minus integration gyroscope bias (GetDeltaBias, imu_preintegration.cpp:151-155)
  double dT;
  double info[81], igrw[9], iarw[9];  // information matrices (not yet scaled by the penalties)
  double pad;
};

// Model of one observation: residual, Huber weight, Jacobians JR, Jt, Jp (2x3 row-major each).
struct ObsModel {
  double r[2], w;
  double JR[6], Jt[6], Jp[6];
  double pad;
};

// Model of one inertial edge: residuals and the Jacobian blocks (9x3 row-major) with respect to
// r1 t1 v1 bg1 ba1 r2 t2 v2.
struct EdgeModel {
  double r[9], rg[3], ra[3];
  double J[8][27];
  // the edge's Gauss-Newton blocks with its penalty: H = J^T (pen Lambda) J over the 24 variables, g = J^T (pen
  // Lambda) r (the random walks and priors are added separately)
  double H[24][24];
  double g[24];
  int32_t active, pad;
};

// Scalars of one LM iteration, written by the reduction kernel and read by the host.
struct IterScalars {
  double cost;    // trial cost (FP64)
  double pred;    // d^T (H + 2 lambda diag H) d
  double maxdp, maxdc;
  int32_t skipped;
  int32_t chol_ok;
  int32_t pad[2];
};

struct Window {
  int32_t P, N, K, F, n, C;  // points, observations, key frames, fixed key frames, free key frames, cameras
  int32_t npairs, nblocks;
  double g[3];
  double scale;  // Huber scale
  double prior_gyro, prior_acc, imu_penalty, acc_rw_penalty, boundary_penalty;
  double lambda;
  // inputs
  const double* xy;      // [N][2]
  const double* info;    // [N][4] row-major
  const int32_t* opid;   // [N] point of each observation
  const int32_t* okid;   // [N] key frame
  const int32_t* ocid;   // [N] camera
  const double* cam;     // [C][12]: R (row-major) and t of cam_from_imu
  const int32_t* point_obs;  // [P+1] into pobs
  const int32_t* pobs;       // [N]
  const int32_t* point_pair; // [P+1]
  const int32_t* pair_obs;   // [npairs][2]: begin, end into pobs
  const int32_t* pair_pose;  // [npairs]
  const int32_t* pair_point; // [npairs]
  const int32_t* pose_obs;   // [n+1] into kobs
  const int32_t* kobs;       // [N]
  const int32_t* pose_pair;  // [n+1] into kpair
  const int32_t* kpair;      // [npairs]
  const int32_t* blk_jk;     // [nblocks][2] (j, k), j >= k
  const int32_t* blk_tri;    // [nblocks+1]
  const int2* tri;           // (pair of j, pair of k)
  const Preint* pre;         // [K]
  // state: [0] current, [1] trial
  double* pts[2];   // [P][3]
  double* pose[2];  // [K][kPS]
  // model and systems
  ObsModel* om;     // [N]
  EdgeModel* em;    // [K-1]
  double* U;        // [15n][15n]
  double* ea;       // [15n]
  double* V;        // [P][9]
  double* eb;       // [P][3]
  double* W;        // [npairs][18]: 3x6 row-major (point rows, pose columns r, t)
  double* Vinv;     // [P][9] damped point-block inverse
  double* rp;       // [P][3] Vinv e_b
  double* Y;        // [npairs][18]: Vinv W
  double* A;        // [15n][15n] reduced system, then its Cholesky factor (lower)
  double* rhs;      // [15n] reduced right-hand side, then the pose step
  double* dp;       // [P][3]
  double* ocost;    // [N] trial cost of each observation (-1: skipped)
  double* ecost;    // [K] inertial terms of each edge, and the bias prior in the last slot
  double* ppred;    // [P] per-point share of the prediction
  double* pmax;     // [P] per-point max |dp|
  double* cpred;    // [15n] per-row share of the pose prediction
  IterScalars* out;
  int32_t* chol_ok;
};

}  // namespace sba_imu_dev
