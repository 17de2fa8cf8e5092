// Drop-in cuvslam::sba::MonoSBASolver (libs/sba/mono_sba_solver.h:90-191) over sba_mono_solve_batch: the window
// collected by addFrame / addTrack is handed over as plain buffers (removed tracks and their observations left
// out), and the solved poses, points, covariance and statistics are kept for the getters.
#include "sba/mono_sba_solver.h"

#include <cstring>

#include "cuvslam_sba.h"
#include "cuvslam_sba_mono.h"

namespace cuvslam::sba {

class MonoSBASolver::Impl {
public:
  sba_handle* h = nullptr;
};

MonoSBASolver::MonoSBASolver(size_t maxFrames, size_t maxTracks, int iterations, float minResidual,
                             float minResidualSpeed)
    : iterations_(-1), max_iterations_(iterations), min_residual_(minResidual),
      min_residual_speed_(minResidualSpeed), impl(new Impl) {
  frames_.reserve(maxFrames);
  tracks_.reserve(maxTracks);
  sba_options o;
  sba_default_options(&o);
  if (sba_create(&o, &impl->h) != SBA_OK) impl->h = nullptr;
}

MonoSBASolver::~MonoSBASolver() {
  if (impl && impl->h) sba_destroy(impl->h);
}

void MonoSBASolver::addFrame(FrameId id, bool constrained, const Isometry3T& a,
                             const std::map<TrackId, bool>& filtered,
                             const std::vector<camera::Observation>& trackCoords) {
  const size_t f = frames_.size();
  map_frames_.emplace(id, f);
  frames_.push_back(FrameIn{id, constrained, a});
  for (const auto& c : trackCoords) {
    const auto it = filtered.find(c.id);
    if (it != filtered.end() && it->second) continue;
    auto t = map_tracks_.find(c.id);
    if (t == map_tracks_.end()) {  // a track seen before addTrack gives its point later
      t = map_tracks_.emplace(c.id, tracks_.size()).first;
      tracks_.push_back(TrackIn{c.id, Vector3T(), false});
    }
    obs_.push_back(ObsIn{static_cast<int32_t>(f), static_cast<int32_t>(t->second), c.xy});
  }
}

void MonoSBASolver::addTrack(TrackId id, const Vector3T& b) {
  auto t = map_tracks_.find(id);
  if (t == map_tracks_.end()) {
    t = map_tracks_.emplace(id, tracks_.size()).first;
    tracks_.push_back(TrackIn{id, b, false});
  } else {
    tracks_[t->second].b = b;
    tracks_[t->second].removed = false;
  }
  ++num_tracks_;
}

size_t MonoSBASolver::numFrames() const { return map_frames_.size(); }
size_t MonoSBASolver::numTracks() const { return num_tracks_; }

bool MonoSBASolver::removeTrack(TrackId id) {
  const auto t = map_tracks_.find(id);
  if (t != map_tracks_.end()) tracks_[t->second].removed = true;
  return true;
}

void MonoSBASolver::resetState() {  // mono_sba_solver.cpp:155-182
  if (failed_) {
    residual_ = -1.f;
    failed_ = false;
  }
  lambda_ = 0.001f;
  iterations_ = -1;
}

void MonoSBASolver::reset() {  // mono_sba_solver.cpp:184-203
  frames_.clear();
  tracks_.clear();
  obs_.clear();
  map_frames_.clear();
  map_tracks_.clear();
  num_tracks_ = 0;
  resetState();
}

int MonoSBASolver::solve() {
  // compact the window: removed tracks and their observations are left out
  std::vector<int32_t> remap(tracks_.size(), -1);
  std::vector<float> pts;
  for (size_t t = 0; t < tracks_.size(); ++t)
    if (!tracks_[t].removed) {
      remap[t] = static_cast<int32_t>(pts.size() / 3);
      for (int d = 0; d < 3; ++d) pts.push_back(tracks_[t].b(d));
    }
  std::vector<int32_t> con(frames_.size()), of, ot;
  std::vector<float> poses(16 * frames_.size()), xy;
  for (size_t f = 0; f < frames_.size(); ++f) {
    con[f] = frames_[f].constrained ? 1 : 0;
    std::memcpy(&poses[16 * f], frames_[f].pose.data(), 64);
  }
  for (const auto& o : obs_) {
    if (remap[o.track] < 0) continue;
    of.push_back(o.frame);
    ot.push_back(remap[o.track]);
    xy.push_back(o.xy(0));
    xy.push_back(o.xy(1));
  }
  sba_mono_window w{static_cast<int32_t>(frames_.size()), static_cast<int32_t>(pts.size() / 3),
                    static_cast<int32_t>(of.size()), con.data(), poses.data(), pts.data(), of.data(), ot.data(),
                    xy.data(), max_iterations_, min_residual_, min_residual_speed_};
  std::vector<float> out_poses(poses.size()), out_pts(pts.size());
  sba_mono_result r{};
  r.frame_poses = out_poses.data();
  r.points = out_pts.data();
  if (!impl->h || sba_mono_solve_batch(impl->h, &w, 1, &r) != SBA_OK) {
    failed_ = true;
    return iterations_ = 0;
  }
  for (size_t f = 0; f < frames_.size(); ++f) std::memcpy(frames_[f].pose.data(), &out_poses[16 * f], 64);
  for (size_t t = 0; t < tracks_.size(); ++t)
    if (remap[t] >= 0)
      for (int d = 0; d < 3; ++d) tracks_[t].b(d) = out_pts[3 * remap[t] + d];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) covariance_(i, j) = r.covariance[6 * i + j];
  residual_ = static_cast<float>(r.residual);
  residual_initial_ = static_cast<float>(r.residual_initial);
  lambda_ = static_cast<float>(r.lambda);
  failed_ = r.failed != 0;
  return iterations_ = r.iterations;
}

float MonoSBASolver::getResidual() const { return residual_; }
float MonoSBASolver::getResidualInitial() const { return residual_initial_; }
float MonoSBASolver::getLambda() const { return lambda_; }
int MonoSBASolver::getIterations() const { return iterations_; }
bool MonoSBASolver::getFailed() const { return failed_; }
Matrix6T MonoSBASolver::getCovariance() const { return covariance_; }

CameraMap MonoSBASolver::getCameras() const {
  CameraMap cameras;
  for (const auto& f : frames_) cameras[f.id] = f.pose;
  return cameras;
}

Tracks3DMap MonoSBASolver::getTracks3d() const {
  Tracks3DMap out;
  for (const auto& t : tracks_)
    if (!t.removed) out[t.id] = t.b;
  return out;
}

}  // namespace cuvslam::sba
