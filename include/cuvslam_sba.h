/*
 * cuvslam_sba.h -- C ABI of the MI355X sliding-window sparse bundle adjustment (SBA).
 *
 * This is the drop-in boundary for cuVSLAM's visual SBA solve:
 *   cuvslam::sba::SchurComplementBundlerGpu::solve(BundleAdjustmentProblem&)
 *     (reference: libs/sba/schur_complement_bundler_gpu.h:29-38,
 *                 libs/sba/schur_complement_bundler_gpu.cpp:115-259)
 * with the in/out problem of libs/sba/bundle_adjustment_problem.h:34-68.
 *
 * Every buffer is a plain pointer with the memory layout of the reference's std::vector<...>::data():
 *   points            Vector3f   -> float[3]   per point
 *   observation_xys   Vector2f   -> float[2]   per observation (normalized image coordinates)
 *   observation_infos Matrix2f   -> float[4]   per observation, column-major 2x2 information matrix
 *   point_ids/pose_ids int32, camera_ids int8
 *   rig_from_world    Isometry3f -> float[16]  per key frame, column-major 4x4
 *   camera_from_rig   Isometry3f -> float[16]  per camera (camera::Rig::camera_from_rig, libs/camera/rig.h:29-43)
 * The last num_fixed_points points and the last num_fixed_key_frames key frames do not move
 * (bundle_adjustment_problem.h:35,52-56).
 *
 * No function throws. Every function returns an sba_status; sba_last_error() describes the last failure.
 * One handle is used by one thread at a time (the reference bundler has the same contract).
 */
#ifndef CUVSLAM_SBA_H_
#define CUVSLAM_SBA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBA_ABI_VERSION 1

typedef enum sba_status {
  SBA_OK = 0,              /* solve ran; the reference bundler returns true */
  SBA_E_EMPTY = 1,         /* no points, poses or observations (schur_complement_bundler_gpu.cpp:121-123) */
  SBA_E_INFEASIBLE = 2,    /* every observation is inside the near plane at the start (gpu.cpp:165-168) */
  SBA_E_DEVICE = 3,        /* a HIP runtime error */
  SBA_E_CAPACITY = 4,      /* a window exceeds what the kernels support */
  SBA_E_INVALID = 5,       /* a bad argument (null pointer, index out of range, ...) */
  SBA_E_NO_PROGRESS = 6    /* inertial SBA: the final cost is not below the initial one; the reference returns
                              false (imu_sba_gpu_fixed_vel.cpp:362) */
} sba_status;

typedef enum sba_loss {
  SBA_LOSS_STUDENT_T = 0,  /* the reference's hard-wired loss: rho(s) = (d+1) log(1+s/d)
                              (schur_complement_bundler_cpu.cpp:39-40, math/robust_cost_function.cpp:89-95) */
  SBA_LOSS_HUBER = 1       /* rho(s) = s/2 for s < d^2, else d sqrt(s) - d^2/2 (robust_cost_function.cpp:62-87) */
} sba_loss;

/* The problem of one window. Mirrors cuvslam::sba::BundleAdjustmentProblem field by field. */
typedef struct sba_problem_view {
  int32_t num_points;
  int32_t num_fixed_points;
  int32_t num_observations;
  int32_t num_poses;
  int32_t num_fixed_key_frames;
  int32_t num_cameras;
  const float* points;            /* [num_points][3] */
  const float* observation_xys;   /* [num_observations][2] */
  const float* observation_infos; /* [num_observations][4], column-major */
  const int32_t* point_ids;       /* [num_observations] */
  const int32_t* pose_ids;        /* [num_observations] */
  const int8_t* camera_ids;       /* [num_observations] */
  const float* rig_from_world;    /* [num_poses][16], column-major */
  const float* camera_from_rig;   /* [num_cameras][16], column-major */
  int32_t max_iterations;         /* BundleAdjustmentProblem::max_iterations (default 20) */
  float robustifier_scale;        /* BundleAdjustmentProblem::robustifier_scale (default 1) */
} sba_problem_view;

/* The solver statistics and final state of one window. */
typedef struct sba_result {
  float* points;                  /* out [num_points][3]; may alias the problem's points */
  float* rig_from_world;          /* out [num_poses][16]; may alias the problem's poses */
  int32_t iterations;             /* BundleAdjustmentProblem::iterations */
  float initial_cost;             /* BundleAdjustmentProblem::initial_cost */
  float last_cost;                /* BundleAdjustmentProblem::last_cost */
  int32_t status;                 /* sba_status of this window */
  double initial_cost_f64;        /* the same costs as evaluated (in FP64) by the solver */
  double last_cost_f64;
  float lambda;                   /* final LM damping */
  int32_t accepted;               /* number of accepted LM steps */
} sba_result;

typedef struct sba_options {
  int32_t device;                 /* HIP device ordinal */
  int32_t loss;                   /* sba_loss; SBA_LOSS_STUDENT_T matches the reference */
  int32_t fixed_iterations;       /* 1: run exactly max_iterations LM iterations (the step-size
                                     convergence exit is off; the zero-cost and infeasible guards stay) */
  int32_t use_graph;              /* 1: capture the LM iterations of a batch in one hipGraph */
  int32_t chunk_pairs;            /* work per Schur workgroup in (point, key frame) pairs; 0 = built-in
                                     heuristic (sba_autotune_chunks picks a value by timing) */
  int32_t pack_threads;           /* host threads that sort and pack one window; 0 = up to 16 for windows of 65536
                                     observations or more, else 1 (the packed streams do not depend on it) */
  int32_t bchunk_pairs;           /* work per back-substitution workgroup in pairs; 0 = chunk_pairs */
  int32_t reserved[5];
} sba_options;

typedef struct sba_handle sba_handle;
typedef struct sba_batch sba_batch;

/* Fills *opt with the defaults (device 0, Student-t loss, reference early exit, graphs on, autotuned tiles). */
void sba_default_options(sba_options* opt);

int sba_create(const sba_options* opt, sba_handle** out);
int sba_destroy(sba_handle* h);
const char* sba_last_error(const sba_handle* h);
int sba_abi_version(void);

/* Solves one window in place, like SchurComplementBundlerGpu::solve (schur_complement_bundler_gpu.cpp:93-100).
 * The return value is the window's status; SBA_OK corresponds to the reference's `true`. */
int sba_solve(sba_handle* h, const sba_problem_view* problem, sba_result* result);

/* sba_solve without any device allocation: no reference counterpart. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The window's
 * streams, solver state and reduced system are carved out of a fixed device scratch array of 64 MiB that the
 * library holds on each GPU it runs on until the process ends (sba_create looks up its address); no hipMalloc,
 * hipFree or other allocating call is made, and the solve is launched eagerly on the handle's stream (what must
 * start at zero is cleared by a fill kernel, not a memset). The host side packs into ordinary host memory.
 * A window that needs more than the scratch holds returns SBA_E_CAPACITY and changes nothing (sba_solve takes any
 * size). Every handle on one GPU shares that GPU's array: calls that overlap on one GPU (several threads or
 * streams) race on it and must be serialized by the caller. Results are those of sba_solve. */
int sba_solve_noalloc(sba_handle* h, const sba_problem_view* problem, sba_result* result);

/* Solves n independent windows of uneven size in one batch: every kernel of an LM iteration covers all
 * windows, and the iterations run as one hipGraph. No reference counterpart (the reference solves one
 * window per call). Returns SBA_OK when the batch ran; per-window statuses are in results[i].status. */
int sba_solve_batch(sba_handle* h, const sba_problem_view* problems, int32_t n, sba_result* results);

/* Device-resident batches, for callers that solve the same packed windows repeatedly (serving, timing):
 * upload and pack once, then run the LM iterations from the uploaded initial state as often as needed. */
int sba_batch_create(sba_handle* h, const sba_problem_view* problems, int32_t n, sba_batch** out);
int sba_batch_run(sba_handle* h, sba_batch* b);       /* asynchronous on the handle's stream */
int sba_batch_sync(sba_handle* h);                     /* waits for the handle's stream */
int sba_batch_read(sba_handle* h, sba_batch* b, sba_result* results);   /* synchronous */
int sba_batch_destroy(sba_handle* h, sba_batch* b);
/* Number of observations, pairs, chunks and windows of a batch; and one eager solve timed with HIP events
 * around every launch (total milliseconds and launches per kernel class, 5 classes:
 * [linearize+Schur, record reduction, factor/solve, back-substitute+cost, decide]). */
int sba_batch_info(sba_handle* h, sba_batch* b, int64_t* num_observations, int64_t* num_pairs,
                   int32_t* num_chunks, int32_t* num_windows);
int sba_batch_kernel_times(sba_handle* h, sba_batch* b, float* ms_per_class, int32_t* launches_per_class);

/* Autotuner of the batched launch's work granularity (pairs per Schur chunk; the back-substitution chunks follow
 * it). Each candidate builds the batch, runs one warm-up solve and times `repeats` solves with HIP events on the
 * handle's stream; the fastest is returned in best_chunk_pairs (set it as sba_options.chunk_pairs of a handle).
 * candidates == NULL or n_candidates == 0 tries the built-in heuristic scaled by 1/4 .. 4; ms_per_candidate
 * [optional, n_candidates] receives the mean time per solve of each given candidate. No reference counterpart
 * (the reference solves one window per call). */
int sba_autotune_chunks(sba_handle* h, const sba_problem_view* problems, int32_t n, const int32_t* candidates,
                        int32_t n_candidates, int32_t repeats, int32_t* best_chunk_pairs, float* ms_per_candidate);

/* The autotuner with a table: looks up the batch's shape class (device architecture, loss, window count, total
 * observations to two significant digits, largest free key-frame count) in the text file `path`; when it is not
 * there, tunes the Schur chunk size (sba_autotune_chunks over its default candidates), then the back-substitution
 * chunk size (1, 2, 4 and 8 times the Schur size), and appends the result. *from_table [optional] is 1 when the
 * sizes came from the file. Set the results as sba_options.chunk_pairs and bchunk_pairs of a handle. No reference
 * counterpart. */
int sba_autotune_table(sba_handle* h, const char* path, const sba_problem_view* problems, int32_t n, int32_t repeats,
                       int32_t* chunk_pairs, int32_t* bchunk_pairs, int32_t* from_table);

/* Landmark winsorizing between two solves (replaces cuvslam::pipelines::winsorize, pipelines/winsorizer.cpp:23-73,
 * as called between the two bundler runs of service_sba.h:317-339), on the solved state of `window` of a batch
 * (run sba_batch_run first). The squared reprojection errors (normalized coordinates) of the observations made
 * from key frame `key_frame` are taken, skipping points inside the near plane; with at least 50 of them, the
 * int(5 n / 100) largest reset their points when above 0.0025^2 (ties at the cut by decreasing point index; the
 * reference orders ties by landmark address). Reset points' observations then carry zero information in the
 * batch, and the window's solved state becomes its initial state, so the next sba_batch_run is the reference's
 * second run: it continues from the first solve without the reset landmarks.
 *   reset_mask [optional, num_points of the window] 1 for reset points; n_reset [optional] their number.
 * SBA_E_CAPACITY when more than 1024 observations share the cut error (nothing is changed then). */
int sba_batch_winsorize(sba_handle* h, sba_batch* b, int32_t window, int32_t key_frame, uint8_t* reset_mask,
                        int32_t* n_reset);

/* One LM step of window 0 of a batch from its current state with damping lambda, for phase-level parity
 * tests (the reference tests its phases the same way: libs/cuda_modules/test/sba_test.cpp).
 * Outputs (any may be null):
 *   S       [n*n] FP64 damped reduced camera system, n = 6 * free key frames (before factorization)
 *   e       [n]   reduced right-hand side
 *   dc      [n]   pose step;  dp [3 * free points] point step
 *   scalars [8]   trial cost, number skipped, hessian term d^T H d, scaling term d^T diag(H) d, max|dp|, max|dc|,
 *                 current cost, factorization succeeded (1) or not (0)
 * Returns SBA_OK, or SBA_E_INFEASIBLE when the factorization failed (S not positive definite). */
int sba_batch_debug_step(sba_handle* h, sba_batch* b, double lambda, double* S, double* e, double* dc,
                         double* dp, double* scalars);

/* The LM iterations of `window` in the last run of a batch, as the device decided them (the first 64 are kept), for
 * parity tests of the damping control (schur_complement_bundler_gpu.cpp:175-253): per iteration the gain ratio
 * (0 when none was formed), the FP64 trial cost (0 when the factorization failed), the damping the step was solved
 * with, and flags (1 accepted, 2 factorization failed, 4 a stop test ended the solve). *n receives the number of
 * iterations written (at most capacity). Any output may be null. */
int sba_batch_trace(sba_handle* h, sba_batch* b, int32_t window, int32_t capacity, double* rho, double* trial_cost,
                    float* lambda, int32_t* flags, int32_t* n);

/* FP64 refinement of the solved state of window 0 of a batch, for the differentiable path: the FP32 LM solve
 * stops at its own rounding floor, and an implicit gradient taken there carries that residual gradient. The FP32
 * state is copied to FP64 (free rotations projected onto SO(3)), then
 *   gauss_newton_steps steps x <- x (-) M^-1 grad F with M = 2 sum rho' J^T L J, the Gauss-Newton system of the
 *     FP32 solution factored once in FP64 (positive definite, so each step descends), then
 *   newton_steps steps x <- x (-) H^-1 grad F with the full FP64 Hessian (the conjugate gradients of
 *     sba_batch_adjoint, truncated at non-positive curvature),
 * where x (-) a is T_j <- Exp(-a_j) T_j, p_i <- p_i - a_i. The refined state stays with the batch and is the state
 * sba_batch_adjoint differentiates at, until the next sba_batch_run; the batch's FP32 state and its
 * sba_batch_read results are unchanged. Zero steps only copy the FP32 state. Outputs (any may be null):
 *   points         [num_points][3] FP64
 *   rig_from_world [num_poses][16] FP64 4x4, column-major (the Eigen layout of Isometry3T::matrix())
 *   grad_norm      |grad F| at the refined state.
 * No reference counterpart. */
int sba_batch_refine(sba_handle* h, sba_batch* b, int32_t gauss_newton_steps, int32_t newton_steps, int32_t max_cg,
                     double tol, double* points, double* rig_from_world, double* grad_norm);

/* Implicit gradient through the solved state of window 0 of a batch (run sba_batch_run first; the solve should
 * have converged, and sba_batch_refine polishes it in FP64). g is dL/dx for a loss L of the solution in the solver's tangent coordinates: a left twist
 * (omega, v) of each free key frame (6 each, stored order), then each free point (3 each). Solves H a = g with
 * H the full FP64 Hessian of the robust cost at the solution (preconditioned conjugate gradients, the
 * preconditioner being the Gauss-Newton system factored once through its reduced camera system), and returns
 *   dxy   [num_observations][2] dL/d observation_xys, in the input observation order
 *   dinfo [num_observations][3] dL/d observation_infos as (L00, L01 = L10 shared, L11)
 *   a_out [optional] the adjoint a.
 * No reference counterpart (the reference has no differentiation). */
int sba_batch_adjoint(sba_handle* h, sba_batch* b, const double* g, double* dxy, double* dinfo, int32_t max_cg,
                      double tol, int32_t* cg_iterations, double* a_out);

/* Factor/solve of one symmetric n x n system (n a multiple of 6, n <= 240) with the reduced-camera-system
 * kernel: A row-major (only its lower triangle is read), b [n] -> x [n]. Returns SBA_OK, or
 * SBA_E_INFEASIBLE when A is not positive definite (the reference's devInfo != 0,
 * libs/cuda_modules/sba.cpp:619-642). For the factorization known-answer tests. */
int sba_debug_factor(sba_handle* h, int32_t n, const double* A, const double* b, double* x);

/* Copies `bytes` between two device buffers `repeats` times with a plain streaming kernel (k_copy_calibrate)
 * and returns the mean ms per copy: the known-size reference that calibrates HBM byte counters. */
int sba_debug_copy(sba_handle* h, int64_t bytes, int32_t repeats, float* ms);

/* One FP32 16x16x4 matrix-core product D = A B (A 16x4, B 4x16, D 16x16, row-major) on one wavefront:
 * checks the lane layout the Schur kernel relies on. */
int sba_debug_mfma_layout(sba_handle* h, const float* A, const float* B, float* D);

/* ---- Synthetic scenes (the benchmark generator; host only, no device needed) -------------------------- */
typedef struct sba_scene_config {
  int32_t num_keyframes;          /* key frames */
  int32_t num_points;             /* 3-D points */
  int32_t sample_observations;    /* 0: a point is observed wherever it projects in-image;
                                     1: each point is observed by U{2..num_keyframes} key frames it projects in */
  int32_t num_cameras;            /* 1 (mono pinhole) */
  int32_t num_fixed_keyframes;    /* the first key frames are fixed (stored last); 1 leaves the scale free */
  double fx, fy, cx, cy;
  int32_t width, height;
  double keyframe_spacing;        /* metres along +x */
  double keyframe_rotation_deg;   /* sigma of each key frame's random rotation */
  double pixel_noise;             /* sigma, pixels */
  double outlier_rate;            /* fraction of observations replaced by a uniform pixel */
  double pose_sigma_m, pose_sigma_deg, point_sigma_m;   /* initial-state perturbation */
  double huber_px;                /* information = diag(fx^2, fy^2) / huber_px^2 */
  double xy_extent, z_min, z_max; /* points: x, y in [-xy_extent, xy_extent], z in [z_min, z_max] */
  uint64_t seed;
} sba_scene_config;

/* The large multi-camera map: a rig of stereo pairs (yaw offsets 360/num_stereo_pairs apart, baseline
 * stereo_baseline_m) on num_keyframes key frames spaced evenly on a closed loop of loop_length_m, inside a corridor
 * of half-width corridor_half_m around the loop (inner and outer walls, floor, ceiling) with a fraction of clutter
 * points in the corridor volume. A camera sees a point that projects in-image within max_range_m; each point keeps
 * a random subset U{2..n} of the n key frames that see it (all of their cameras that see it). Points are stored in
 * order of their angle around the loop. */
typedef struct sba_loop_config {
  int32_t num_keyframes, num_points, num_stereo_pairs, num_fixed_keyframes;
  double loop_length_m, stereo_baseline_m, corridor_half_m, floor_y_m, ceiling_y_m, clutter_fraction, max_range_m;
  double fx, fy, cx, cy;
  int32_t width, height;
  double keyframe_rotation_deg, pixel_noise, outlier_rate, pose_sigma_m, pose_sigma_deg, point_sigma_m, huber_px;
  uint64_t seed;
} sba_loop_config;

typedef struct sba_scene sba_scene;

void sba_scene_default_config(sba_scene_config* cfg);
void sba_loop_default_config(sba_loop_config* cfg);
int sba_scene_generate_loop(const sba_loop_config* cfg, sba_scene** out);
int sba_scene_generate(const sba_scene_config* cfg, sba_scene** out);
/* Fills a problem view that points into the scene's own storage (valid until sba_scene_free). */
int sba_scene_view(const sba_scene* s, sba_problem_view* view);
/* Ground-truth points [num_points][3] and rig_from_world [num_poses][16] in the stored order. */
int sba_scene_ground_truth(const sba_scene* s, const float** points, const float** rig_from_world);
void sba_scene_free(sba_scene* s);

#ifdef __cplusplus
}
#endif

#endif /* CUVSLAM_SBA_H_ */
