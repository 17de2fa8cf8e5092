// The part of cuVSLAM's map that the SBA service reads and writes: key frames and landmarks shared with the
// tracking thread, and the recent submap the service turns into a BundleAdjustmentProblem.
//
// Mirrors the members the service uses of libs/map/map.h (KeyFrame::get_pose/set_pose, Landmark::get_pose/
// set_pose/reset, UnifiedMap::SubMap::consecutive_keyframes and ::landmark_and_obs) and camera::Observation
// (libs/camera/observation.h: cam_id, id, xy, xy_info). Poses are guarded by a mutex each, so the tracking thread
// may triangulate (set a pose) or reset landmarks while the service marshals a problem.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <vector>

#include "sba/bundle_adjustment_problem.h"

namespace cuvslam {
using TrackId = uint64_t;
using FrameId = int64_t;
}  // namespace cuvslam

namespace cuvslam::camera {

// One 2-D observation of a track (libs/camera/observation.h): normalized image coordinates and their information
// matrix, in camera cam_id of the rig.
struct Observation {
  int32_t cam_id = 0;
  TrackId id = 0;
  Vector2T xy;
  Matrix2T xy_info = Matrix2T::Identity();
  Observation() = default;
  Observation(int32_t cam, TrackId track, const Vector2T& p, const Matrix2T& info)
      : cam_id(cam), id(track), xy(p), xy_info(info) {}
};

}  // namespace cuvslam::camera

namespace cuvslam::map {

class KeyFrame {
public:
  KeyFrame(const Isometry3T& rig_from_world, FrameId id) : id_(id), pose_(rig_from_world) {}
  FrameId id() const { return id_; }
  Isometry3T get_pose() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return pose_;
  }
  void set_pose(const Isometry3T& rig_from_world) {
    std::lock_guard<std::mutex> lock(mutex_);
    pose_ = rig_from_world;
  }

private:
  FrameId id_;
  mutable std::mutex mutex_;
  Isometry3T pose_;
};

// A landmark has a world position once triangulated; reset() removes it (the winsorizer does this to exclude a
// landmark from the next solve).
class Landmark {
public:
  explicit Landmark(TrackId id) : id_(id) {}
  Landmark(TrackId id, const Vector3T& world) : id_(id), pose_(world) {}
  TrackId id() const { return id_; }
  std::optional<Vector3T> get_pose() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return pose_;
  }
  void set_pose(const Vector3T& world) {
    std::lock_guard<std::mutex> lock(mutex_);
    pose_ = world;
  }
  void reset() {
    std::lock_guard<std::mutex> lock(mutex_);
    pose_.reset();
  }

private:
  TrackId id_;
  mutable std::mutex mutex_;
  std::optional<Vector3T> pose_;
};

using KeyframePtr = std::shared_ptr<KeyFrame>;
using LandmarkPtr = std::shared_ptr<Landmark>;

// The most recent key frames of the map, oldest first, with the landmarks each one observes.
struct SubMap {
  struct KeyframeEntry {
    KeyframePtr keyframe;
  };
  struct LandmarkAndObserv {
    LandmarkPtr landmark;
    std::vector<camera::Observation> observations;  // one per camera of the rig that sees the landmark
  };
  std::vector<KeyframeEntry> consecutive_keyframes;
  std::vector<std::vector<LandmarkAndObserv>> landmark_and_obs;  // parallel to consecutive_keyframes
};

// Where the service takes its submap from (UnifiedMap::get_recent_submap in the reference).
class SubMapSource {
public:
  virtual ~SubMapSource() = default;
  virtual SubMap get_recent_submap(int32_t num_keyframes) = 0;
};

}  // namespace cuvslam::map
