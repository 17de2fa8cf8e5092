// Device-side data layout of a packed SBA batch (shared by the HIP kernels and the host runtime).
//
// A batch is any number of independent windows packed back to back in structure-of-arrays streams:
//   observations  sorted by (point, key frame, camera) inside each window:
//                   obs_xy   float2  normalized image coordinates
//                   obs_info float4  (L00, (L01+L10)/2, L11, camera id as int bits)
//   pairs         one per distinct (point, key frame) of a window, i.e. one block W_ij of the
//                 Hessian's off-diagonal part; a point's pairs are sorted by key frame, so its pairs with
//                 free key frames come first:
//                   pair_obs  int32 first observation of the pair (pair_obs[p+1] ends it)
//                   pair_pose int32 key frame index inside the window
//                   pair_rec  int4  (first observation, observation end, key frame, point index in its Schur tile)
//                   pair_xy, pair_info  the pair's first observation, indexed by pair
//   points          point_pair int32 first pair of the point (point_pair[i+1] ends it)
//                   point_nfree int32 number of the point's pairs with free key frames
//                   pts[2]   float4 (x, y, z, 0), double-buffered: current and trial state
//                   paux     3 x float4: damped inverse V*^-1 (6 upper entries), V*^-1 e_b (3),
//                            diag(V) (3), written by the Schur kernel for the back-substitution
//   key frames      poses[2] 3 x float4 rows of [R | t] of rig_from_world, double-buffered
//   cameras         cams     3 x float4 rows of [R | t] of camera_from_rig
//   reduced system  part     one record per chunk: its Schur products sum_i Z_i Z_i^T (Z = W G, G G^T the
//                            damped point-block pseudo-inverse) as upper 16x16 tiles in matrix-core lane
//                            order, then U (21 upper entries per free key frame), e (6), diag U (6)
//                   sred     the window's records summed in chunk order (deterministic)
//                   A        FP64 Cholesky workspace of mapped windows (more than kMaxFreePoses free key frames)
//                   dc       FP32 pose step (6 per free key frame)
// Windows carry their offsets into every stream, so one launch covers all windows.
#pragma once

#include <stdint.h>

namespace sba_dev {

constexpr int kBlock = 256;        // threads of the per-tile kernels
constexpr int kTilePairs = 256;    // max (point, key frame) pairs per tile (one per thread)
constexpr int kTilePoints = 16;    // max points per tile: 48 Schur columns, 12 matrix-core k-steps
constexpr int kMaxFreePoses = 40;  // free key frames per window (reduced system order <= 240)
constexpr int kMaxPoses = 1024;  // key frames per window
constexpr int kMaxCameras = 32;    // camera::Rig::kMaxCameras (libs/camera/rig.h:30)
constexpr int kNT = 15;            // 16x16 tiles per side of the largest reduced system (240 / 16)
constexpr int kZcLd = 240;         // leading dimension of the staged Z columns (bank-conflict free)
constexpr int kKt = 48;            // Schur columns per tile (3 per point)
constexpr int kFactorThreads = 512;
constexpr int kPanel = 8;          // Cholesky panel width
constexpr float kFltEps = 1.1920928955078125e-7f;  // FLT_EPSILON, common/types.h:203-206

struct WindowDesc {
  int32_t np, npf, nc, ncf, ncam, no;
  int32_t point_off, pair_off, obs_off, pose_off, cam_off;
  int32_t nt;        // 16x16 tiles per side of the reduced system, ceil(6 ncf / 16)
  int32_t rec;       // floats of one Schur record: upper tiles (256 each), U (21 ncf), e (6 ncf), diag U (6 ncf)
  int32_t nchunks;   // Schur records (chunks) of the window
  int64_t part_off;  // floats into part: the window's first chunk record
  int64_t sred_off;  // floats into sred: the window's reduced record
  int64_t a_off;     // doubles into A (mapped windows)
  int32_t dc_off;    // floats into dc
  int32_t max_iterations;
  float robust;
  // windows with more than kMaxFreePoses free key frames ("mapped"): each chunk stages its own <= kMaxFreePoses
  // free key frames; the chunk records are gathered, in chunk order, into a dense FP64 system in A
  int32_t mapped;
  int32_t chunk_begin;  // the window's first Schur chunk
  int32_t gfirst_off;   // ints into gfirst: ncf * ncf + 1 offsets, one per key-frame block (a, b), row-major
  int32_t bchunk_begin, nbchunks;  // the window's back-substitution chunks (and their CostPart slots)
};

struct WindowState {
  double current_cost, initial_cost, last_cost;
  double scal_pose;  // d_c^T diag(U) d_c of the last step (the factor kernel)
  float lambda, maxdc;
  int32_t chol_ok, done, status, iterations, accepted, cur;
};

// One back-substitution chunk's share of a window's trial cost and predicted-reduction terms. Each chunk writes its
// own slot; the decision kernel sums a window's slots in chunk order, so the totals do not depend on scheduling.
struct CostPart {
  double cost, hess, scal;  // FP64 trial cost, d^T H d and d^T diag(V) d of the chunk's observations and points
  int32_t skipped;          // observations inside the near plane
  float maxdp;              // max |dp| over the chunk's points
};

// One LM iteration of a window as the decision kernel saw it (the first kTrace iterations are kept).
constexpr int kTrace = 64;
struct TraceEntry {
  double rho;         // gain ratio (0 when no ratio was formed)
  double trial_cost;  // FP64 cost of the trial state (0 when the factorization failed)
  float lambda;       // damping the iteration's step was solved with
  int32_t flags;      // 1 accepted, 2 factorization failed, 4 a stop test ended the solve
  int32_t pad[2];
};

struct Tile {
  int32_t window, point_begin, point_end, pair_begin, pair_end;
  // A point of a mapped window seen from more than kMaxFreePoses free key frames has no place in the ordinary
  // Schur tiles: its free key frames are split into groups of at most kMaxFreePoses / 2, and for every two groups
  // a tile of its own stages their union, over a copy of all the point's pairs (pair_begin .. pair_end) in which
  // the pairs of the other groups count as fixed. pair_shift = pair_begin - point_pair[point]; virt = 1 | the
  // chunk keeps the blocks inside the first group << 1 | inside the second << 2 | size of the first group << 8
  // (0 for ordinary tiles).
  int32_t pair_shift, virt, pad;
};

struct Chunk {
  int32_t tile_begin, tile_end;
  int64_t part_off;  // floats into part: this chunk's record
  int32_t nloc;      // free key frames staged by the chunk (the window's free key frames unless mapped)
  int32_t nt;        // 16x16 tiles per side of the chunk's staged system, ceil(6 nloc / 16)
  int32_t map_off;   // ints into chunk_map: the chunk's free key frames in increasing order (mapped windows)
  int32_t window;
};

struct Batch {
  int32_t n_windows, n_chunks, n_points, n_poses, n_cams;
  int32_t loss, fixed_iterations, max_rec;
  int32_t max_nc;  // key frames of the largest window (sizes the back-substitution's shared memory)
  const WindowDesc* win;
  WindowState* st;
  const Tile* tiles;     // Schur tiles (<= kTilePairs pairs, <= kTilePoints points)
  const Chunk* chunks;
  const Tile* btiles;    // back-substitution tiles (<= kTilePairs pairs, <= kBlock points)
  const Chunk* bchunks;
  int32_t n_bchunks, red_blocks;  // red_blocks: workgroups of the record reduction
  const int32_t* red_first;       // per window (+1): its first reduction workgroup
  const float2* obs_xy;
  const float4* obs_info;
  const int32_t* pair_obs;
  const int32_t* pair_pose;
  const int4* pair_rec;      // per pair: (first observation, end of its observations, key frame | point index
                             // inside its back-substitution tile << 16, point index inside its Schur tile |
                             // local free key frame index << 8, 255 when fixed)
  const float2* pair_xy;     // per pair: its first observation (the obs streams themselves when every pair has
  const float4* pair_info;   // exactly one observation)
  const int32_t* point_pair;
  const int32_t* point_nfree;
  float4* pts0;
  float4* pts1;
  const float4* pts_init;
  float4* paux;      // 3 per point
  float4* poses0;    // 3 per pose
  float4* poses1;
  const float4* poses_init;
  const float4* cams;
  const int32_t* chunk_map;  // local -> window free key frame, per chunk of a mapped window
  // mapped windows: the contributions to key-frame block (a, b), a <= b, of the reduced system are
  // gcon[gfirst[W.gfirst_off + a ncf + b] .. gfirst[... + 1]), in chunk order, each (chunk, la | lb << 8 |
  // nt << 16 | nloc << 24, part_off low and high 32 bits) with la, lb the chunk-local indices of a and b and
  // nt, nloc, part_off those of the chunk
  const int32_t* gfirst;
  const int4* gcon;
  float* part;       // per-chunk Schur records
  float* sred;       // per-window reduced records
  double* A;
  float* dc;
  double* dc64;
  float4* out_pts;
  float4* out_poses;
  float4* dbg_dp;    // optional: the point step of the last back-substitution (phase-level tests)
  TraceEntry* trace;  // optional: kTrace entries per window
  CostPart* bpart;    // one slot per back-substitution chunk
  // shared-memory layout of the Schur kernel (offsets in 4-byte words)
  int32_t lds_pose, lds_cam, lds_acc, lds_ent, lds_ec, lds_w, lds_pts, lds_zc, lds_words;
};

// Orthonormal polar factor of a near-rotation (Newton iteration X <- (X + X^-T) / 2, Higham 1986); for a
// matrix with positive determinant it is U V^T of its SVD, which common/rotation_utils.h:36-50 returns.
__device__ __forceinline__ void polar3(double X[3][3]) {
  for (int it = 0; it < 3; ++it) {
    const double c00 = X[1][1] * X[2][2] - X[1][2] * X[2][1];
    const double c01 = X[1][2] * X[2][0] - X[1][0] * X[2][2];
    const double c02 = X[1][0] * X[2][1] - X[1][1] * X[2][0];
    const double c10 = X[0][2] * X[2][1] - X[0][1] * X[2][2];
    const double c11 = X[0][0] * X[2][2] - X[0][2] * X[2][0];
    const double c12 = X[0][1] * X[2][0] - X[0][0] * X[2][1];
    const double c20 = X[0][1] * X[1][2] - X[0][2] * X[1][1];
    const double c21 = X[0][2] * X[1][0] - X[0][0] * X[1][2];
    const double c22 = X[0][0] * X[1][1] - X[0][1] * X[1][0];
    const double det = X[0][0] * c00 + X[0][1] * c01 + X[0][2] * c02;
    const double id = 1.0 / det;  // X^-T = cofactor(X) / det
    const double C[3][3] = {{c00, c01, c02}, {c10, c11, c12}, {c20, c21, c22}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) X[i][j] = 0.5 * (X[i][j] + C[i][j] * id);
  }
}


// Exp of a twist (omega, v) in FP64, the formula of math/twist.cpp:23-41 (theta clamped at FLT_EPSILON).
__device__ __forceinline__ void twist_exp64(const double* tw, double R[3][3], double t[3]) {
  const double w0 = tw[0], w1 = tw[1], w2 = tw[2];
  const double theta = fmax(sqrt(w0 * w0 + w1 * w1 + w2 * w2), static_cast<double>(kFltEps));
  const double c = cos(theta), st = sin(theta) / theta;
  const double n[3] = {w0 / theta, w1 / theta, w2 / theta};
  const double ws[3] = {w0 * st, w1 * st, w2 * st};
  const double K[3][3] = {{c, -ws[2], ws[1]}, {ws[2], c, -ws[0]}, {-ws[1], ws[0], c}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = K[i][j] + (1.0 - c) * n[i] * n[j];
  const double* v = tw + 3;
  const double nxv[3] = {n[1] * v[2] - n[2] * v[1], n[2] * v[0] - n[0] * v[2], n[0] * v[1] - n[1] * v[0]};
  const double nnv[3] = {n[1] * nxv[2] - n[2] * nxv[1], n[2] * nxv[0] - n[0] * nxv[2], n[0] * nxv[1] - n[1] * nxv[0]};
  for (int i = 0; i < 3; ++i) t[i] = v[i] + ((1.0 - c) / theta * nxv[i] + (1.0 - st) * nnv[i]);
}

// Trial key frames of window W from the FP64 pose step dc: Exp(dc_j) * T_j, rotation re-orthonormalized, rounded
// to FP32 into the non-current pose buffer; fixed key frames are copied.
__device__ __forceinline__ void write_trial_poses(const Batch& b, const WindowDesc& W, int cur, const double* dc,
                                                  int tid, int nthreads) {
  const int ncf = W.ncf;
  const float4* src = (cur ? b.poses1 : b.poses0) + 3 * W.pose_off;
  float4* dst = (cur ? b.poses0 : b.poses1) + 3 * W.pose_off;
  for (int j = tid; j < W.nc; j += nthreads) {
    const float4 r0 = src[3 * j], r1 = src[3 * j + 1], r2 = src[3 * j + 2];
    if (j >= ncf) {
      dst[3 * j] = r0;
      dst[3 * j + 1] = r1;
      dst[3 * j + 2] = r2;
      continue;
    }
    double E[3][3], et[3];
    twist_exp64(dc + 6 * j, E, et);
    const double R[3][3] = {{r0.x, r0.y, r0.z}, {r1.x, r1.y, r1.z}, {r2.x, r2.y, r2.z}};
    const double t[3] = {r0.w, r1.w, r2.w};
    double X[3][3], tn[3];
    for (int a = 0; a < 3; ++a) {
      for (int c = 0; c < 3; ++c) X[a][c] = E[a][0] * R[0][c] + E[a][1] * R[1][c] + E[a][2] * R[2][c];
      tn[a] = E[a][0] * t[0] + E[a][1] * t[1] + E[a][2] * t[2] + et[a];
    }
    polar3(X);
    dst[3 * j] = make_float4(static_cast<float>(X[0][0]), static_cast<float>(X[0][1]), static_cast<float>(X[0][2]),
                             static_cast<float>(tn[0]));
    dst[3 * j + 1] = make_float4(static_cast<float>(X[1][0]), static_cast<float>(X[1][1]), static_cast<float>(X[1][2]),
                                 static_cast<float>(tn[1]));
    dst[3 * j + 2] = make_float4(static_cast<float>(X[2][0]), static_cast<float>(X[2][1]), static_cast<float>(X[2][2]),
                                 static_cast<float>(tn[2]));
  }
}

}  // namespace sba_dev
