// The SBA service around the bundler: problem assembly from the map's recent submap, the worker thread that runs it
// when triggered, and the winsorizer between two solves. Header-only; any type with
// `bool solve(sba::BundleAdjustmentProblem&)` is a bundler (the MI355X SchurComplementBundlerGpu, or a test double).
//
// Mirrors libs/pipelines/service_sba.h:41-180,301-348 (SbaServiceBase, run_sba, SbaService),
// libs/pipelines/service_sba.cpp:22-33 (CalcNumFixedKeyframes), libs/pipelines/winsorizer.cpp:28-70 (winsorize),
// libs/map/service.cpp:26-95 (the worker loop) and libs/sba/sba_config.h:27-46 (the settings).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "sba/bundle_adjustment_problem.h"
#include "sba/map.h"

namespace cuvslam::sba {

// libs/sba/sba_config.h:27-46 (the members the visual SBA service reads)
struct Settings {
  bool async = true;
  int32_t num_sba_frames = 7;        // key frames in a window
  int32_t num_fixed_sba_frames = 3;  // of which fixed (the oldest)
  int32_t num_sba_iterations = 7;    // LM iterations per solve
  float robustifier_scale = 5e-1f;
  bool use_sba_winsorizer = false;
};

}  // namespace cuvslam::sba

namespace cuvslam::pipelines {

// How many of the window's key frames to fix: none for fewer than two key frames (nothing to solve), one when
// there are not enough to fix the requested number and keep one free (service_sba.cpp:22-33).
inline int CalcNumFixedKeyframes(size_t map_size, size_t num_fixed_key_frames) {
  if (map_size < 2) return 0;
  if (map_size < num_fixed_key_frames + 1) return 1;
  return static_cast<int>(num_fixed_key_frames);
}

struct KeyframeLandmarkObs {
  map::KeyframePtr keyframe;
  map::LandmarkPtr landmark;
  camera::Observation observation;
};

// Landmark winsorizing (winsorizer.cpp:28-70): of the observations made from the last key frame, those whose
// landmark has a position in front of the camera (depth > 0.1, near_plane.h:33-38) are ranked by squared
// reprojection error in normalized coordinates; with at least 50 of them, the landmarks of the int(5 n / 100)
// largest errors above 0.0025^2 are reset, which removes them from the next problem. Returns the number reset.
inline int winsorize(const camera::Rig& rig, const std::vector<KeyframeLandmarkObs>& observations_from_last_keyframe) {
  constexpr float kPercent = 5.f;
  constexpr float kMinSize = 0.0025f;
  constexpr size_t kSafe = 50;
  std::vector<std::pair<float, map::LandmarkPtr>> errors;
  errors.reserve(observations_from_last_keyframe.size());
  for (const auto& o : observations_from_last_keyframe) {
    const auto point = o.landmark->get_pose();
    if (!point) continue;
    const Isometry3T& c = rig.camera_from_rig[o.observation.cam_id];
    const Isometry3T r = o.keyframe->get_pose();
    float pr[3], pc[3];
    for (int i = 0; i < 3; ++i)
      pr[i] = r(i, 0) * point->x() + r(i, 1) * point->y() + r(i, 2) * point->z() + r(i, 3);
    for (int i = 0; i < 3; ++i) pc[i] = c(i, 0) * pr[0] + c(i, 1) * pr[1] + c(i, 2) * pr[2] + c(i, 3);
    if (!(pc[2] > 0.1f)) continue;
    const float dx = pc[0] / pc[2] - o.observation.xy.x(), dy = pc[1] / pc[2] - o.observation.xy.y();
    errors.emplace_back(dx * dx + dy * dy, o.landmark);
  }
  const size_t n = errors.size();
  if (n < kSafe) return 0;
  std::sort(errors.begin(), errors.end(),
            [](const auto& a, const auto& b) { return a.first < b.first || (a.first == b.first && a.second < b.second); });
  const int to_delete = static_cast<int>(kPercent * static_cast<float>(n) / 100.f);
  int reset = 0;
  for (int j = 0; j < to_delete; ++j) {
    const auto& [err, landmark] = errors[n - 1 - static_cast<size_t>(j)];
    if (err > kMinSize * kMinSize) {
      landmark->reset();
      ++reset;
    }
  }
  return reset;
}

// Builds the window's problem from the submap, solves it and writes the solution back (service_sba.h:72-180):
//   * the oldest CalcNumFixedKeyframes(...) key frames are fixed and stored last; the others keep their order;
//   * every landmark with a position is a point, numbered in order of first appearance (key frames oldest first,
//     each key frame's landmarks in order); landmarks without a position are skipped, also one that gains its
//     position from another thread between the two passes below;
//   * observations are ordered by key frame, then landmark, then camera;
//   * poses and points are written back only when the bundler returns true.
// Returns the bundler's result (false, with nothing solved, when there are fewer than two key frames).
template <class Bundler>
bool run_sba(const map::SubMap& recent_map, const camera::Rig& rig, const sba::Settings& settings, Bundler& bundler) {
  sba::BundleAdjustmentProblem problem;
  problem.rig = rig;
  problem.num_fixed_points = 0;
  problem.num_fixed_key_frames =
      CalcNumFixedKeyframes(recent_map.consecutive_keyframes.size(), static_cast<size_t>(settings.num_fixed_sba_frames));
  if (problem.num_fixed_key_frames < 1) return false;

  // first pass: the points (landmarks with a position, each once)
  std::unordered_map<const map::Landmark*, std::pair<map::LandmarkPtr, int32_t>> point_of;
  size_t total_observations = 0;
  for (const auto& landmarks : recent_map.landmark_and_obs)
    for (const auto& lo : landmarks) {
      const auto p = lo.landmark->get_pose();
      if (!p) continue;
      if (point_of.find(lo.landmark.get()) == point_of.end()) {
        point_of.emplace(lo.landmark.get(), std::make_pair(lo.landmark, static_cast<int32_t>(problem.points.size())));
        problem.points.push_back(*p);
      }
      total_observations += lo.observations.size();
    }

  // pose ids: free key frames first, the fixed (oldest) ones last
  const size_t nk = recent_map.consecutive_keyframes.size();
  const size_t nfix = static_cast<size_t>(problem.num_fixed_key_frames);
  std::vector<int32_t> pose_of(nk);
  for (size_t i = nfix; i < nk; ++i) {
    pose_of[i] = static_cast<int32_t>(problem.rig_from_world.size());
    problem.rig_from_world.push_back(recent_map.consecutive_keyframes[i].keyframe->get_pose());
  }
  for (size_t i = 0; i < nfix; ++i) {
    pose_of[i] = static_cast<int32_t>(problem.rig_from_world.size());
    problem.rig_from_world.push_back(recent_map.consecutive_keyframes[i].keyframe->get_pose());
  }

  // second pass: the observations of the indexed points
  problem.observation_xys.reserve(total_observations);
  problem.observation_infos.reserve(total_observations);
  problem.point_ids.reserve(total_observations);
  problem.pose_ids.reserve(total_observations);
  problem.camera_ids.reserve(total_observations);
  for (size_t i = 0; i < nk && i < recent_map.landmark_and_obs.size(); ++i) {
    for (const auto& lo : recent_map.landmark_and_obs[i]) {
      const auto it = point_of.find(lo.landmark.get());
      if (it == point_of.end()) continue;  // no position during the first pass
      for (const auto& o : lo.observations) {
        problem.observation_xys.push_back(o.xy);
        problem.observation_infos.push_back(o.xy_info);
        problem.point_ids.push_back(it->second.second);
        problem.pose_ids.push_back(pose_of[i]);
        problem.camera_ids.push_back(static_cast<int8_t>(o.cam_id));
      }
    }
  }

  problem.max_iterations = settings.num_sba_iterations;
  problem.robustifier_scale = settings.robustifier_scale;
  const bool res = bundler.solve(problem);
  if (res) {
    for (size_t i = 0; i < nk; ++i) recent_map.consecutive_keyframes[i].keyframe->set_pose(problem.rig_from_world[pose_of[i]]);
    for (const auto& [raw, entry] : point_of) entry.first->set_pose(problem.points[entry.second]);
  }
  return res;
}

// The worker of a map service (libs/map/service.cpp:26-95): trigger() hands over the settings of the next run and
// wakes the worker; in synchronous mode it runs the task on the caller's thread. The settings are snapshotted under
// the lock just before a task starts, so a trigger during a running task is neither lost nor seen half-written
// (service_sba.h:41-69).
class SbaServiceBase {
public:
  SbaServiceBase(map::SubMapSource& map, std::string name, bool async) : map_(map), async_(async), name_(std::move(name)) {}
  virtual ~SbaServiceBase() { stop(); }
  SbaServiceBase(const SbaServiceBase&) = delete;
  SbaServiceBase& operator=(const SbaServiceBase&) = delete;

  void trigger(const sba::Settings& s) {
    if (!async_) {
      pending_ = s;
      current_ = pending_;
      service_task();
      finished();
      return;
    }
    {
      std::lock_guard<std::mutex> lock(mutex_);
      pending_ = s;
      inputs_ready_ = true;
    }
    cv_.notify_one();
  }

  // Blocks until `count` tasks have finished since the service started (for callers that need the result).
  void wait_for_tasks(uint64_t count) {
    std::unique_lock<std::mutex> lock(mutex_);
    done_cv_.wait(lock, [&] { return tasks_done_ >= count; });
  }
  uint64_t tasks_done() {
    std::lock_guard<std::mutex> lock(mutex_);
    return tasks_done_;
  }
  const std::string& name() const { return name_; }

protected:
  virtual void service_task() = 0;

  // Starts the worker; derived constructors call it once their members exist.
  void start() {
    if (!async_ || job_.joinable()) return;
    stop_ = false;
    job_ = std::thread([this] { run(); });
  }
  // Waits for the running task and ends the worker; derived destructors call it before their members go away.
  void stop() {
    if (!async_) return;
    {
      std::lock_guard<std::mutex> lock(mutex_);
      stop_ = true;
    }
    cv_.notify_one();
    if (job_.joinable()) job_.join();
  }

  map::SubMapSource& map_;
  sba::Settings pending_;  // written by trigger() under the lock
  sba::Settings current_;  // the running task's snapshot

private:
  void run() {
    while (true) {
      {
        std::unique_lock<std::mutex> lock(mutex_);
        cv_.wait(lock, [&] { return inputs_ready_ || stop_; });
        if (stop_) return;
        inputs_ready_ = false;  // a trigger during the task sets it again: that run is not lost
        current_ = pending_;
      }
      service_task();
      finished();
    }
  }
  void finished() {
    {
      std::lock_guard<std::mutex> lock(mutex_);
      ++tasks_done_;
    }
    done_cv_.notify_all();
  }

  const bool async_;
  const std::string name_;
  bool inputs_ready_ = false;
  bool stop_ = false;
  uint64_t tasks_done_ = 0;
  std::mutex mutex_;
  std::condition_variable cv_, done_cv_;
  std::thread job_;
};

// The visual SBA service (service_sba.h:301-348): each task solves the recent submap; with the winsorizer on, it
// then resets the landmarks of the worst observations from the newest key frame and solves again without them.
template <class Bundler>
class SbaService : public SbaServiceBase {
public:
  template <class... BundlerArgs>
  SbaService(const sba::Settings& settings, const camera::Rig& rig, map::SubMapSource& map, BundlerArgs&&... args)
      : SbaServiceBase(map, "cuvslam:sba", settings.async), rig_(rig), bundler_(std::forward<BundlerArgs>(args)...) {
    pending_ = settings;
    current_ = settings;
    start();
  }
  ~SbaService() override { stop(); }

  Bundler& bundler() { return bundler_; }
  // Results of the last task: whether each solve returned true, and the landmarks the winsorizer reset.
  bool last_first_solve() const { return first_ok_; }
  bool last_second_solve() const { return second_ok_; }
  int last_winsorized() const { return winsorized_; }

protected:
  void service_task() final {
    const map::SubMap recent_map = map_.get_recent_submap(current_.num_sba_frames);
    first_ok_ = run_sba(recent_map, rig_, current_, bundler_);
    second_ok_ = false;
    winsorized_ = 0;
    if (current_.use_sba_winsorizer && !recent_map.consecutive_keyframes.empty() && !recent_map.landmark_and_obs.empty()) {
      const map::KeyframePtr last_kf = recent_map.consecutive_keyframes.back().keyframe;
      std::vector<KeyframeLandmarkObs> observations;
      for (const auto& lo : recent_map.landmark_and_obs.back())
        for (const auto& o : lo.observations) observations.push_back({last_kf, lo.landmark, o});
      winsorized_ = winsorize(rig_, observations);
      second_ok_ = run_sba(recent_map, rig_, current_, bundler_);
    }
  }

private:
  camera::Rig rig_;
  Bundler bundler_;
  bool first_ok_ = false, second_ok_ = false;
  int winsorized_ = 0;
};

}  // namespace cuvslam::pipelines
