/*
 * cuvslam_sba_imu.h -- C ABI of the MI355X inertial sliding-window bundle adjustment.
 *
 * This is the drop-in boundary for cuVSLAM's visual-inertial SBA solve:
 *   cuvslam::sba_imu::IMUBundlerGpuFixedVel::solve(ImuBAProblem&)
 *     (reference: libs/imu/imu_sba_gpu.h:42-47, libs/imu/imu_sba_gpu_fixed_vel.cpp:75-363)
 * with the in/out problem of libs/imu/imu_sba_problem.h:30-91 and the IMU preintegration of
 * libs/imu/imu_preintegration.h:33-108.
 *
 * Every key frame carries a 15-dof state: rotation and translation of w_from_imu (perturbed on the right:
 * R <- R Exp(dr), t <- t + R dt), velocity, gyroscope bias and accelerometer bias. Key frame i also carries the
 * preintegration of the IMU samples from key frame i to key frame i+1, which ties the two states with a 9-dof
 * inertial residual and two 3-dof bias random-walk residuals. Observations are the same pinhole reprojections
 * as the visual SBA, through camera_from_rig * rig_from_imu * imu_from_world, with a Huber loss.
 *
 * Buffers have the memory layout of the reference's std::vector<...>::data(): Isometry3f = float[16]
 * column-major 4x4, Matrix3f = float[9] column-major, Matrix2f = float[4] column-major, Matrix<float,9,9> =
 * float[81] column-major. The first num_fixed_key_frames key frames do not move (imu_sba_problem.h:67-69).
 *
 * No function throws; every function returns an sba_status (cuvslam_sba.h), and sba_last_error() of the
 * handle describes the last failure.
 */
#ifndef CUVSLAM_SBA_IMU_H_
#define CUVSLAM_SBA_IMU_H_

#include <stddef.h>
#include <stdint.h>

#include "cuvslam_sba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One IMU sample: common/imu_measurement.h (time in ns, specific force in m/s^2, angular rate in rad/s). */
typedef struct sba_imu_measurement {
  int64_t time_ns;
  float linear_acceleration[3];
  float angular_velocity[3];
} sba_imu_measurement;

/* imu::ImuCalibration (common/imu_calibration.h:29-60): the noise densities and random walks of the Kalibr
 * model; the discrete covariances are density^2 * frequency and walk^2 / frequency (imu_calibration.cpp:35-47). */
typedef struct sba_imu_calibration {
  float rig_from_imu[16];            /* column-major 4x4 */
  float gyroscope_noise_density;     /* rad / (s sqrt(Hz)) */
  float gyroscope_random_walk;       /* rad / (s^2 sqrt(Hz)) */
  float accelerometer_noise_density; /* m / (s^2 sqrt(Hz)) */
  float accelerometer_random_walk;   /* m / (s^3 sqrt(Hz)) */
  float frequency;                   /* Hz */
} sba_imu_calibration;

/* The state of an IMUPreintegration (imu_preintegration.h:35-108): the integrated deltas, their bias Jacobians
 * and covariances, the biases it was integrated at and the updated biases (SetNewBias). Its samples are
 * measurements[first_measurement .. first_measurement + num_measurements) of the problem. */
typedef struct sba_imu_preintegration {
  float cov[81];           /* [rotation, velocity, translation] covariance, column-major */
  float acc_rw_cov[9];     /* accumulated accelerometer random-walk covariance */
  float gyro_rw_cov[9];    /* accumulated gyroscope random-walk covariance */
  float dR[9];
  float dV[3];
  float dP[3];
  float JRg[9];
  float JVg[9];
  float JVa[9];
  float JPg[9];
  float JPa[9];
  float dT_s;
  float gyro_bias[3];         /* the biases of the integration */
  float acc_bias[3];
  float gyro_bias_updated[3]; /* SetNewBias */
  float acc_bias_updated[3];
  int32_t first_measurement;
  int32_t num_measurements;
} sba_imu_preintegration;

/* The problem of one inertial window: cuvslam::sba_imu::ImuBAProblem field by field. */
typedef struct sba_imu_problem_view {
  int32_t num_points;
  int32_t num_observations;
  int32_t num_poses;
  int32_t num_fixed_key_frames;
  int32_t num_cameras;
  int32_t num_measurements;
  const float* points;                           /* [num_points][3] */
  const float* observation_xys;                  /* [num_observations][2] */
  const float* observation_infos;                /* [num_observations][4], column-major */
  const int32_t* point_ids;                      /* [num_observations] */
  const int32_t* pose_ids;                       /* [num_observations] */
  const int8_t* camera_ids;                      /* [num_observations] */
  const float* camera_from_rig;                  /* [num_cameras][16] */
  const float* w_from_imu;                       /* [num_poses][16] */
  const float* velocity;                         /* [num_poses][3] */
  const float* gyro_bias;                        /* [num_poses][3] */
  const float* acc_bias;                         /* [num_poses][3] */
  const sba_imu_preintegration* preintegration;  /* [num_poses]: key frame i to i+1 */
  const sba_imu_measurement* measurements;       /* [num_measurements] */
  float gravity[3];                              /* world-frame gravity acceleration */
  int32_t max_iterations;                        /* ImuBAProblem defaults: 7 */
  float robustifier_scale;                       /* 1 */
  float robustifier_scale_pose;                  /* 10 (not used by the GPU bundler) */
  float prior_gyro;                              /* 0 */
  float prior_acc;                               /* 0 */
  float imu_penalty;                             /* 1 */
  float acc_rw_penalty;                          /* -1: use imu_penalty */
  float boundary_imu_penalty;                    /* 1e-2 */
  float reintegration_thresh;                    /* 1e-4 */
} sba_imu_problem_view;

/* The final state and the statistics of one inertial window. Every output may alias the problem's input. */
typedef struct sba_imu_result {
  float* points;                          /* out [num_points][3] */
  float* w_from_imu;                      /* out [num_poses][16] */
  float* velocity;                        /* out [num_poses][3] */
  float* gyro_bias;                       /* out [num_poses][3] */
  float* acc_bias;                        /* out [num_poses][3] */
  sba_imu_preintegration* preintegration; /* out [num_poses] (re-integrated where the solve re-integrates); may be
                                             null */
  int32_t status;                         /* SBA_OK when the reference would return true */
  int32_t iterations;                     /* ImuBAProblem::iterations */
  int32_t accepted;
  float initial_cost;                     /* ImuBAProblem::initial_cost */
  float lambda;                           /* final LM damping */
  int32_t pad;
  double initial_cost_f64;
  double last_cost_f64;                   /* the cost of the returned state */
} sba_imu_result;

/* Solves one inertial window in place (IMUBundlerGpuFixedVel::solve, imu_sba_gpu_fixed_vel.cpp:75-363): LM with
 * lambda_0 = 1e2, x0.125 / x5, gain-ratio thresholds 0.25 / 0.75; a failed factorization or a non-finite trial
 * cost raises lambda; the preintegration of a free key frame is re-integrated at its new biases when they moved
 * by more than reintegration_thresh. Returns SBA_OK when the final cost is below the initial one (the reference's
 * true), SBA_E_NO_PROGRESS when it is not, SBA_E_EMPTY without an inertial edge, points or observations, and
 * SBA_E_INFEASIBLE when the initial cost is not finite. The calibration supplies rig_from_imu and the noise model of
 * re-integrations. */
int sba_imu_solve(sba_handle* h, const sba_imu_calibration* calib, const sba_imu_problem_view* problem,
                  sba_imu_result* result);

/* IMUPreintegration(gyro_bias, acc_bias) followed by IntegrateNewMeasurement of n samples
 * (imu_preintegration.cpp:26-150): the first sample's step is 1/frequency, later ones the time difference.
 * out->first_measurement is set to 0 and out->num_measurements to n. */
int sba_imu_integrate(const sba_imu_calibration* calib, const sba_imu_measurement* samples, int32_t n,
                      const float gyro_bias[3], const float acc_bias[3], sba_imu_preintegration* out);

/* InfoMatrix, InfoGyroRWMatrix, InfoAccRWMatrix (imu_preintegration.cpp:190-240): the inverses of the three
 * covariances (zero for an empty preintegration or a covariance that is not positive definite), column-major. */
int sba_imu_information(const sba_imu_preintegration* p, float info[81], float info_gyro_rw[9],
                        float info_acc_rw[9]);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* CUVSLAM_SBA_IMU_H_ */
