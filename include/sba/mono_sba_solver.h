// Drop-in MonoSBASolver (libs/sba/mono_sba_solver.h:90-191) backed by the MI355X batched mono solve
// (cuvslam_sba_mono.h): the same namespace, type names and methods, so the monocular pipeline's caller
// (pipelines/sba_mono_wrapper.cpp) compiles unchanged. addFrame / addTrack collect the window; solve() runs it on the
// GPU (one workgroup, the whole LM loop on the device) from the current poses and points; the getters return its
// results. Each solve starts from identity track information matrices and lambda 1e-3, as a first solve does.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "sba/bundle_adjustment_problem.h"
#include "sba/small_matrix.h"

namespace cuvslam {

using FrameId = size_t;                           // common/frame_id.h:25
using TrackId = size_t;                           // common/track_id.h:24
using CameraId = uint8_t;                         // common/camera_id.h:24
using CameraMap = std::map<FrameId, Isometry3T>;  // common/isometry_utils.h:66
using Tracks3DMap = std::map<TrackId, Vector3T>;  // common/vector_3t.h:34

namespace camera {
struct Observation {  // camera/observation.h:29-40
  Observation() = default;
  Observation(CameraId cam_id_, TrackId id_, const Vector2T& xy_, const Matrix2T& xy_info_)
      : cam_id(cam_id_), id(id_), xy(xy_), xy_info(xy_info_) {}
  CameraId cam_id = 0;
  TrackId id = 0;
  Vector2T xy;
  Matrix2T xy_info;
};
}  // namespace camera

namespace sba {

class MonoSBASolver {
public:
  MonoSBASolver(size_t maxFrames, size_t maxTracks, int iterations = 1000, float minResidual = 0.001f,
                float minResidualSpeed = 0.001f);
  ~MonoSBASolver();
  MonoSBASolver(const MonoSBASolver&) = delete;
  MonoSBASolver& operator=(const MonoSBASolver&) = delete;

  // a: world_from_camera of the frame; observations of tracks marked filtered are left out (mono_sba_solver.cpp:33-103)
  void addFrame(FrameId id, bool constrained, const Isometry3T& a, const std::map<TrackId, bool>& filtered,
                const std::vector<camera::Observation>& trackCoords);
  void addTrack(TrackId id, const Vector3T& b);
  size_t numFrames() const;
  size_t numTracks() const;
  bool removeTrack(TrackId id);
  void resetState();
  void reset();

  // Runs the solve; returns the number of LM iterations (mono_sba_solver.cpp:205-301).
  int solve();

  float getResidual() const;
  float getResidualInitial() const;
  float getLambda() const;
  int getIterations() const;
  CameraMap getCameras() const;
  Tracks3DMap getTracks3d() const;
  bool getFailed() const;
  Matrix6T getCovariance() const;

private:
  struct FrameIn {
    FrameId id;
    bool constrained;
    Isometry3T pose;  // world_from_camera
  };
  struct TrackIn {
    TrackId id;
    Vector3T b;
    bool removed;
  };
  struct ObsIn {
    int32_t frame, track;
    Vector2T xy;
  };
  int iterations_, max_iterations_;
  float min_residual_, min_residual_speed_;
  float residual_ = -1.f, residual_initial_ = 0.f, lambda_ = 0.001f;
  bool failed_ = false;
  size_t num_tracks_ = 0;
  std::vector<FrameIn> frames_;
  std::vector<TrackIn> tracks_;
  std::vector<ObsIn> obs_;
  std::unordered_map<FrameId, size_t> map_frames_;
  std::unordered_map<TrackId, size_t> map_tracks_;
  Matrix6T covariance_;
  class Impl;
  std::unique_ptr<Impl> impl;
};

}  // namespace sba
}  // namespace cuvslam
