// CPU tests of the SBA service (include/sba/service_sba.h) with a recording bundler: the cases of
// libs/pipelines/test/service_sba_test.cpp (landmarks without a position are skipped; problem construction
// tolerates triangulation from another thread) and of the service's contract (fixed key frames, problem order,
// write-back, the worker thread, the winsorizer). Prints one line per case and exits non-zero on a failure.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "sba/service_sba.h"

using namespace cuvslam;
using cuvslam::map::KeyFrame;
using cuvslam::map::Landmark;
using cuvslam::map::LandmarkPtr;

static int g_failures = 0;
#define CHECK(cond)                                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
      ++g_failures;                                                            \
    }                                                                          \
  } while (0)

// Records the problem it is given and returns `result` (service_sba_test.cpp:39-47).
struct RecordingSbaBundler {
  bool result = true;
  int calls = 0;
  sba::BundleAdjustmentProblem last;
  bool solve(sba::BundleAdjustmentProblem& problem) {
    ++calls;
    last = problem;
    for (auto& p : problem.points) p.x() += 1.f;  // a "solution" that differs from the input
    for (auto& T : problem.rig_from_world) T(0, 3) += 1.f;
    return result;
  }
};

constexpr size_t kNumKeyframes = 4;
constexpr size_t kLandmarksPerKeyframe = 600;

struct Fixture {
  map::SubMap submap;
  std::vector<LandmarkPtr> landmarks;
};

// Each key frame observes its own landmarks once; has_pose(i) decides whether landmark i is triangulated
// (service_sba_test.cpp:62-81).
template <class HasPose>
Fixture MakeSubMap(HasPose has_pose, size_t num_keyframes = kNumKeyframes) {
  Fixture f;
  f.submap.landmark_and_obs.resize(num_keyframes);
  for (size_t k = 0; k < num_keyframes; ++k) {
    Isometry3T T;
    T(0, 3) = static_cast<float>(k);
    f.submap.consecutive_keyframes.push_back({std::make_shared<KeyFrame>(T, static_cast<FrameId>(k))});
    for (size_t j = 0; j < kLandmarksPerKeyframe; ++j) {
      const auto id = static_cast<TrackId>(f.landmarks.size());
      auto l = has_pose(id) ? std::make_shared<Landmark>(id, Vector3T(0, 0, 1)) : std::make_shared<Landmark>(id);
      map::SubMap::LandmarkAndObserv lo{l, {camera::Observation(0, id, Vector2T(0, 0), Matrix2T::Identity())}};
      f.submap.landmark_and_obs[k].push_back(lo);
      f.landmarks.push_back(l);
    }
  }
  return f;
}

bool ConsistentProblem(const sba::BundleAdjustmentProblem& p) {
  for (int32_t id : p.point_ids)
    if (id < 0 || static_cast<size_t>(id) >= p.points.size()) return false;
  for (int32_t id : p.pose_ids)
    if (id < 0 || static_cast<size_t>(id) >= p.rig_from_world.size()) return false;
  return p.point_ids.size() == p.observation_xys.size() && p.pose_ids.size() == p.observation_xys.size();
}

void SkipsLandmarksWithoutPose() {
  std::printf("SkipsLandmarksWithoutPose\n");
  auto f = MakeSubMap([](TrackId id) { return id % 2 == 0; });
  RecordingSbaBundler bundler;
  pipelines::run_sba(f.submap, camera::Rig{}, sba::Settings{}, bundler);
  CHECK(bundler.last.points.size() == f.landmarks.size() / 2);
  CHECK(bundler.last.observation_xys.size() == f.landmarks.size() / 2);
  CHECK(ConsistentProblem(bundler.last));
  for (const auto& l : f.landmarks) CHECK(l->get_pose().has_value() == (l->id() % 2 == 0));
}

// Triangulates from another thread while the problem is built (service_sba_test.cpp:90-123): every pose is cleared,
// then set again in reverse order, until the triangulator is destroyed.
class ConcurrentTriangulator {
public:
  explicit ConcurrentTriangulator(const std::vector<LandmarkPtr>& landmarks)
      : thread_([this, &landmarks] {
          running_.store(true, std::memory_order_release);
          while (!stop_.load(std::memory_order_acquire)) {
            for (const auto& l : landmarks) l->reset();
            for (auto it = landmarks.rbegin(); it != landmarks.rend(); ++it) (*it)->set_pose(Vector3T(0, 0, 1));
          }
        }) {
    while (!running_.load(std::memory_order_acquire)) {
    }
  }
  ~ConcurrentTriangulator() {
    stop_.store(true, std::memory_order_release);
    thread_.join();
  }

private:
  std::atomic<bool> running_{false};
  std::atomic<bool> stop_{false};
  std::thread thread_;
};

void ToleratesConcurrentTriangulation(int trials) {
  std::printf("ToleratesConcurrentTriangulation (%d trials)\n", trials);
  for (int t = 0; t < trials; ++t) {
    auto f = MakeSubMap([](TrackId) { return false; });
    RecordingSbaBundler bundler;
    {
      ConcurrentTriangulator tri(f.landmarks);
      pipelines::run_sba(f.submap, camera::Rig{}, sba::Settings{}, bundler);
    }
    if (bundler.calls) CHECK(ConsistentProblem(bundler.last));
  }
}

void FixedKeyframeCount() {
  std::printf("FixedKeyframeCount\n");
  CHECK(pipelines::CalcNumFixedKeyframes(0, 3) == 0);
  CHECK(pipelines::CalcNumFixedKeyframes(1, 3) == 0);
  CHECK(pipelines::CalcNumFixedKeyframes(2, 3) == 1);
  CHECK(pipelines::CalcNumFixedKeyframes(3, 3) == 1);
  CHECK(pipelines::CalcNumFixedKeyframes(4, 3) == 3);
  CHECK(pipelines::CalcNumFixedKeyframes(7, 3) == 3);
  // one key frame: nothing to solve, the bundler is not called
  auto f = MakeSubMap([](TrackId) { return true; }, 1);
  RecordingSbaBundler bundler;
  CHECK(!pipelines::run_sba(f.submap, camera::Rig{}, sba::Settings{}, bundler));
  CHECK(bundler.calls == 0);
}

void ProblemOrderAndWriteBack() {
  std::printf("ProblemOrderAndWriteBack\n");
  // 5 key frames; landmark j of key frame k is also seen by key frame k + 1 (two cameras there)
  const size_t nk = 5;
  map::SubMap sm;
  sm.landmark_and_obs.resize(nk);
  std::vector<LandmarkPtr> lms;
  for (size_t k = 0; k < nk; ++k) {
    Isometry3T T;
    T(1, 3) = static_cast<float>(k);
    sm.consecutive_keyframes.push_back({std::make_shared<KeyFrame>(T, static_cast<FrameId>(k))});
  }
  for (size_t k = 0; k < nk; ++k)
    for (int j = 0; j < 3; ++j) {
      auto l = std::make_shared<Landmark>(lms.size(), Vector3T(static_cast<float>(k), static_cast<float>(j), 5.f));
      lms.push_back(l);
      sm.landmark_and_obs[k].push_back({l, {camera::Observation(0, l->id(), Vector2T(0.1f * j, 0.f), Matrix2T::Identity())}});
      if (k + 1 < nk)
        sm.landmark_and_obs[k + 1].push_back({l, {camera::Observation(0, l->id(), Vector2T(0.f, 0.1f * j), Matrix2T::Identity()),
                                                  camera::Observation(1, l->id(), Vector2T(0.2f, 0.1f * j), Matrix2T::Identity())}});
    }
  sba::Settings s;
  s.num_fixed_sba_frames = 2;
  s.num_sba_iterations = 11;
  s.robustifier_scale = 0.25f;
  RecordingSbaBundler bundler;
  CHECK(pipelines::run_sba(sm, camera::Rig{}, s, bundler));
  const auto& p = bundler.last;
  CHECK(p.num_fixed_key_frames == 2 && p.num_fixed_points == 0);
  CHECK(p.max_iterations == 11 && p.robustifier_scale == 0.25f);
  CHECK(p.rig_from_world.size() == nk);
  // free key frames 2, 3, 4 first (pose ids 0, 1, 2), then the fixed oldest 0, 1 (pose ids 3, 4)
  const int expect_pose_of_kf[nk] = {3, 4, 0, 1, 2};
  for (size_t k = 0; k < nk; ++k) CHECK(p.rig_from_world[expect_pose_of_kf[k]](1, 3) == static_cast<float>(k));
  // points in order of first appearance: key frame 0's three, then key frame 1's new three, ...
  CHECK(p.points.size() == lms.size());
  for (size_t i = 0; i < lms.size(); ++i) CHECK(p.points[i].x() == lms[i]->get_pose()->x() - 1.f);
  // observations by key frame, then landmark, then camera
  size_t o = 0;
  bool order_ok = true;
  for (size_t k = 0; k < nk; ++k)
    for (const auto& lo : sm.landmark_and_obs[k])
      for (const auto& ob : lo.observations) {
        order_ok = order_ok && p.pose_ids[o] == expect_pose_of_kf[k] && p.point_ids[o] == static_cast<int32_t>(lo.landmark->id()) &&
                   p.camera_ids[o] == ob.cam_id && p.observation_xys[o].x() == ob.xy.x() && p.observation_xys[o].y() == ob.xy.y();
        ++o;
      }
  CHECK(order_ok && o == p.observation_xys.size());
  // the solution (shifted by the recording bundler) was written back
  CHECK(lms[0]->get_pose()->x() == 1.f);
  CHECK(sm.consecutive_keyframes[0].keyframe->get_pose()(0, 3) == 1.f);
  // a bundler that returns false changes nothing in the map
  RecordingSbaBundler failing;
  failing.result = false;
  CHECK(!pipelines::run_sba(sm, camera::Rig{}, s, failing));
  CHECK(lms[0]->get_pose()->x() == 1.f);
}

// A fixed submap source that counts its calls.
class StaticSource : public map::SubMapSource {
public:
  explicit StaticSource(map::SubMap sm) : sm_(std::move(sm)) {}
  map::SubMap get_recent_submap(int32_t n) override {
    ++calls;
    last_n = n;
    return sm_;
  }
  std::atomic<int> calls{0};
  std::atomic<int> last_n{0};

private:
  map::SubMap sm_;
};

void ServiceRunsOnItsWorker() {
  std::printf("ServiceRunsOnItsWorker\n");
  auto f = MakeSubMap([](TrackId) { return true; });
  StaticSource src(f.submap);
  sba::Settings s;
  s.async = true;
  pipelines::SbaService<RecordingSbaBundler> svc(s, camera::Rig{}, src);
  sba::Settings t = s;
  t.num_sba_frames = 4;
  t.num_sba_iterations = 5;
  svc.trigger(t);
  svc.wait_for_tasks(1);
  CHECK(src.calls == 1 && src.last_n == 4);
  CHECK(svc.bundler().calls == 1 && svc.bundler().last.max_iterations == 5);
  CHECK(svc.last_first_solve());
  // synchronous mode runs on the caller's thread
  sba::Settings sync = s;
  sync.async = false;
  StaticSource src2(f.submap);
  pipelines::SbaService<RecordingSbaBundler> svc2(sync, camera::Rig{}, src2);
  svc2.trigger(sync);
  CHECK(src2.calls == 1 && svc2.tasks_done() == 1);
}

// Exact observations from the last key frame plus a few planted large errors: the planted landmarks are reset, and
// the second solve goes without them.
void WinsorizerResetsTheWorstLandmarks() {
  std::printf("WinsorizerResetsTheWorstLandmarks\n");
  map::SubMap sm;
  sm.landmark_and_obs.resize(2);
  std::vector<LandmarkPtr> lms;
  for (int k = 0; k < 2; ++k) sm.consecutive_keyframes.push_back({std::make_shared<KeyFrame>(Isometry3T(), k)});
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  const int n = 200;
  for (int j = 0; j < n; ++j) {
    const Vector3T p(u(rng), u(rng), 5.f + u(rng));
    auto l = std::make_shared<Landmark>(j, p);
    lms.push_back(l);
    Vector2T xy(p.x() / p.z(), p.y() / p.z());
    if (j % 20 == 3) xy.x() += 0.05f;  // 10 planted errors of 0.05 (normalized)
    sm.landmark_and_obs[0].push_back({l, {camera::Observation(0, j, Vector2T(p.x() / p.z(), p.y() / p.z()), Matrix2T::Identity())}});
    sm.landmark_and_obs[1].push_back({l, {camera::Observation(0, j, xy, Matrix2T::Identity())}});
  }
  struct Identity {  // returns true without changing the state
    int calls = 0;
    size_t points_seen = 0;
    bool solve(sba::BundleAdjustmentProblem& p) {
      ++calls;
      points_seen = p.points.size();
      return true;
    }
  };
  StaticSource src(sm);
  sba::Settings s;
  s.async = false;
  s.use_sba_winsorizer = true;
  s.num_fixed_sba_frames = 1;
  pipelines::SbaService<Identity> svc(s, camera::Rig{}, src);
  svc.trigger(s);
  // int(5 * 200 / 100) = 10 largest errors, all above 0.0025^2: exactly the planted ones
  CHECK(svc.last_winsorized() == 10);
  int reset = 0;
  for (int j = 0; j < n; ++j) {
    const bool planted = j % 20 == 3;
    CHECK(lms[j]->get_pose().has_value() != planted);
    reset += !lms[j]->get_pose().has_value();
  }
  CHECK(reset == 10);
  CHECK(svc.bundler().calls == 2 && svc.bundler().points_seen == static_cast<size_t>(n - 10));
  // fewer than 50 observations: nothing is reset
  std::vector<pipelines::KeyframeLandmarkObs> few;
  for (int j = 0; j < 49; ++j)
    few.push_back({sm.consecutive_keyframes[1].keyframe, lms[j + 1], sm.landmark_and_obs[1][j + 1].observations[0]});
  CHECK(pipelines::winsorize(camera::Rig{}, few) == 0);
}

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 200;
  SkipsLandmarksWithoutPose();
  ToleratesConcurrentTriangulation(trials);
  FixedKeyframeCount();
  ProblemOrderAndWriteBack();
  ServiceRunsOnItsWorker();
  WinsorizerResetsTheWorstLandmarks();
  std::printf("%s (%d failures)\n", g_failures ? "FAILED" : "OK", g_failures);
  return g_failures ? 1 : 0;
}
