/*
 * cuvslam_sba_mono.h -- C ABI of the MI355X monocular SBA solver.
 *
 * The MI355X path of cuVSLAM's MonoSBASolver (libs/sba/mono_sba_solver.h:90-191, mono_sba_solver.cpp:205-301):
 * Hartley & Zisserman's sparse bundle adjustment (2nd ed., Algorithm A6.4) of a monocular window, with camera poses
 * parameterized by the twist of camera_from_world, the Huber loss (threshold 0.5) on per-track information matrices
 * estimated from the track's residuals, multiplicative damping, a x10 / 10 damping schedule and the mean robust
 * residual as the step test, and the covariance of the last free frame's pose. Windows are solved in batches, one
 * workgroup per window with the whole LM loop on the device (the reference solves one window per call on the CPU).
 *
 * Buffers: frame_poses are world_from_camera Isometry3f (float[16], column-major), as MonoSBASolver::addFrame takes
 * them and getCameras returns them; observations are normalized image coordinates.
 */
#ifndef CUVSLAM_SBA_MONO_H_
#define CUVSLAM_SBA_MONO_H_

#include <stddef.h>
#include <stdint.h>

#include "cuvslam_sba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One window: the frames (addFrame), tracks (addTrack) and observations (the trackCoords of each frame) of a
 * MonoSBASolver, and its constructor's solver settings. */
typedef struct sba_mono_window {
  int32_t num_frames;
  int32_t num_tracks;
  int32_t num_obs;
  const int32_t* constrained; /* [num_frames] 1: the frame does not move */
  const float* frame_poses;   /* [num_frames][16] world_from_camera */
  const float* points;        /* [num_tracks][3] */
  const int32_t* obs_frame;   /* [num_obs] */
  const int32_t* obs_track;   /* [num_obs] */
  const float* obs_xy;        /* [num_obs][2] */
  int32_t max_iterations;     /* MonoSBASolver default 1000 */
  float min_residual;         /* 0.001 */
  float min_residual_speed;   /* 0.001 */
} sba_mono_window;

/* The solved window: getCameras / getTracks3d / getCovariance / getResidual / getResidualInitial / getLambda /
 * getIterations / getFailed, and whether any observation could be used. */
typedef struct sba_mono_result {
  float* frame_poses;        /* out [num_frames][16] world_from_camera; may alias the window's */
  float* points;             /* out [num_tracks][3]; may alias the window's */
  float covariance[36];      /* row-major 6x6 (rotation, translation) of the last free frame */
  double residual;
  double residual_initial;
  double lambda;
  int32_t iterations;
  int32_t failed;            /* lambda passed 1e7 or the iterations ran out */
  int32_t feasible;          /* 0: no observation usable (every track behind a camera or degenerate) */
  int32_t status;            /* SBA_OK, or SBA_E_INFEASIBLE when not feasible */
} sba_mono_result;

/* Solves n independent windows in one launch. Returns SBA_OK when the batch ran; each result's status says whether
 * its window was feasible. Windows of more than 64 free frames return SBA_E_CAPACITY. */
int sba_mono_solve_batch(sba_handle* h, const sba_mono_window* windows, int32_t n, sba_mono_result* results);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* CUVSLAM_SBA_MONO_H_ */
