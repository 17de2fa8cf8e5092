// The SBA service (include/sba/service_sba.h) with the MI355X bundler, end to end on the GPU: a map of the
// reference's production window (7 key frames of which 3 fixed, 7 LM iterations) is built from a synthetic scene;
// run_sba assembles the problem from it, SchurComplementBundlerGpu solves it and the solution is written back.
// Checked:
//   * the solve succeeds and lowers the cost;
//   * the map holds exactly the solved state (the problem assembled again from the updated map equals the
//     bundler's output, bit for bit);
//   * the same assembled problem solved by a second bundler gives the same bits;
//   * the asynchronous SbaService (worker thread, trigger, wait) leaves a copy of the map in the same state;
//   * with the winsorizer on, the service's second solve runs and succeeds.
// Prints one line per case and "OK (0 failures)" at the end; exits non-zero on a failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "cuvslam_sba.h"
#include "sba/schur_complement_bundler_gpu.h"
#include "sba/service_sba.h"

using namespace cuvslam;

static int g_failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++g_failures;                                                   \
    }                                                                 \
  } while (0)

// Forwards to the GPU bundler and keeps copies of the problem it was given and of the solved problem.
struct TeeBundler {
  sba::SchurComplementBundlerGpu gpu{0};  // Student-t, the reference's loss
  sba::BundleAdjustmentProblem in, out;
  int calls = 0;
  bool solve(sba::BundleAdjustmentProblem& p) {
    ++calls;
    in = p;
    const bool ok = gpu.solve(p);
    out = p;
    return ok;
  }
};

// Records the problem without solving (returns false: nothing is written back).
struct AssembleOnly {
  sba::BundleAdjustmentProblem last;
  bool solve(sba::BundleAdjustmentProblem& p) {
    last = p;
    return false;
  }
};

struct Scene {
  std::vector<std::vector<float>> poses;  // scene order: free key frames, then the fixed ones
  std::vector<float> points;
  std::vector<float> xy, info;
  std::vector<int32_t> pid, qid;
  std::vector<int8_t> cid;
  int nfix = 0, ncam = 0;
  std::vector<float> camera_from_rig;
};

static Scene make_scene(uint64_t seed) {
  sba_scene_config cfg;
  sba_scene_default_config(&cfg);
  cfg.seed = seed;
  cfg.num_points = 2000;
  cfg.num_keyframes = 7;
  cfg.num_fixed_keyframes = 3;
  sba_scene* sc = nullptr;
  if (sba_scene_generate(&cfg, &sc) != SBA_OK) std::exit(2);
  sba_problem_view v;
  sba_scene_view(sc, &v);
  Scene s;
  for (int k = 0; k < v.num_poses; ++k) s.poses.emplace_back(v.rig_from_world + 16 * k, v.rig_from_world + 16 * k + 16);
  s.points.assign(v.points, v.points + 3 * v.num_points);
  s.xy.assign(v.observation_xys, v.observation_xys + 2 * v.num_observations);
  s.info.assign(v.observation_infos, v.observation_infos + 4 * v.num_observations);
  s.pid.assign(v.point_ids, v.point_ids + v.num_observations);
  s.qid.assign(v.pose_ids, v.pose_ids + v.num_observations);
  s.cid.assign(v.camera_ids, v.camera_ids + v.num_observations);
  s.nfix = v.num_fixed_key_frames;
  s.ncam = v.num_cameras;
  s.camera_from_rig.assign(v.camera_from_rig, v.camera_from_rig + 16 * v.num_cameras);
  sba_scene_free(sc);
  return s;
}

// The map of a scene: key frames oldest first (the scene's fixed key frames are the oldest), each key frame's
// landmarks in point order with one observation per camera.
struct Map : map::SubMapSource {
  map::SubMap submap;
  std::vector<map::LandmarkPtr> landmarks;
  map::SubMap get_recent_submap(int32_t) override { return submap; }
};

static std::unique_ptr<Map> make_map(const Scene& s) {
  auto m = std::make_unique<Map>();
  const int nk = static_cast<int>(s.poses.size()), ncf = nk - s.nfix;
  std::vector<int> scene_pose(nk);  // submap index -> scene pose
  for (int i = 0; i < nk; ++i) scene_pose[i] = i < s.nfix ? ncf + i : i - s.nfix;
  const int np = static_cast<int>(s.points.size() / 3);
  for (int i = 0; i < np; ++i) {
    Vector3T p;
    for (int d = 0; d < 3; ++d) p(d) = s.points[3 * i + d];
    m->landmarks.push_back(std::make_shared<map::Landmark>(static_cast<TrackId>(i), p));
  }
  m->submap.landmark_and_obs.resize(nk);
  for (int i = 0; i < nk; ++i) {
    Isometry3T T;
    std::memcpy(T.data(), s.poses[scene_pose[i]].data(), sizeof(float) * 16);
    m->submap.consecutive_keyframes.push_back({std::make_shared<map::KeyFrame>(T, static_cast<FrameId>(i))});
    // this key frame's observations grouped by point, in point order
    std::vector<std::vector<int>> by_point(np);
    for (size_t o = 0; o < s.pid.size(); ++o)
      if (s.qid[o] == scene_pose[i]) by_point[s.pid[o]].push_back(static_cast<int>(o));
    for (int pnt = 0; pnt < np; ++pnt) {
      if (by_point[pnt].empty()) continue;
      map::SubMap::LandmarkAndObserv lo;
      lo.landmark = m->landmarks[pnt];
      for (int o : by_point[pnt]) {
        Vector2T xy;
        xy(0) = s.xy[2 * o];
        xy(1) = s.xy[2 * o + 1];
        Matrix2T info;
        std::memcpy(info.data(), &s.info[4 * o], sizeof(float) * 4);
        lo.observations.emplace_back(s.cid[o], static_cast<TrackId>(pnt), xy, info);
      }
      m->submap.landmark_and_obs[i].push_back(lo);
    }
  }
  return m;
}

static camera::Rig make_rig(const Scene& s) {
  camera::Rig rig;
  rig.num_cameras = s.ncam;
  for (int c = 0; c < s.ncam; ++c)
    std::memcpy(rig.camera_from_rig[c].data(), &s.camera_from_rig[16 * c], sizeof(float) * 16);
  return rig;
}

static bool same_state(const sba::BundleAdjustmentProblem& a, const sba::BundleAdjustmentProblem& b) {
  return a.points.size() == b.points.size() && a.rig_from_world.size() == b.rig_from_world.size() &&
         std::memcmp(a.points.data(), b.points.data(), sizeof(float) * 3 * a.points.size()) == 0 &&
         std::memcmp(a.rig_from_world.data(), b.rig_from_world.data(), sizeof(float) * 16 * a.rig_from_world.size()) == 0;
}

static sba::BundleAdjustmentProblem assemble(const Map& m, const camera::Rig& rig, const sba::Settings& settings) {
  AssembleOnly rec;
  pipelines::run_sba(m.submap, rig, settings, rec);
  return rec.last;
}

int main() {
  const Scene scene = make_scene(7);
  const camera::Rig rig = make_rig(scene);
  sba::Settings settings;  // the reference's defaults: 7 key frames, 3 fixed, 7 iterations
  settings.async = false;

  std::printf("run_sba with the GPU bundler: solve, write-back\n");
  auto mapA = make_map(scene);
  TeeBundler tee;
  const bool ok = pipelines::run_sba(mapA->submap, rig, settings, tee);
  CHECK(ok);
  CHECK(tee.calls == 1);
  CHECK(tee.in.num_fixed_key_frames == 3);
  CHECK(tee.out.last_cost < tee.out.initial_cost);
  CHECK(same_state(assemble(*mapA, rig, settings), tee.out));
  std::printf("  observations %zu, iterations %d, cost %.6g -> %.6g\n", tee.in.point_ids.size(), tee.out.iterations,
              tee.out.initial_cost, tee.out.last_cost);

  std::printf("the assembled problem solved by a second bundler gives the same bits\n");
  {
    sba::BundleAdjustmentProblem again = tee.in;
    sba::SchurComplementBundlerGpu second(0);
    CHECK(second.solve(again));
    CHECK(same_state(again, tee.out));
    CHECK(again.iterations == tee.out.iterations);
  }

  std::printf("the asynchronous service leaves its map in the same state\n");
  {
    auto mapB = make_map(scene);
    sba::Settings s = settings;
    s.async = true;
    pipelines::SbaService<TeeBundler> service(s, rig, *mapB);
    service.trigger(s);
    service.wait_for_tasks(1);
    CHECK(service.last_first_solve());
    CHECK(same_state(assemble(*mapB, rig, settings), tee.out));
  }

  std::printf("with the winsorizer, the second solve runs\n");
  {
    auto mapC = make_map(scene);
    sba::Settings s = settings;
    s.use_sba_winsorizer = true;
    pipelines::SbaService<TeeBundler> service(s, rig, *mapC);
    service.trigger(s);
    service.wait_for_tasks(1);
    CHECK(service.last_first_solve());
    CHECK(service.last_second_solve());
    CHECK(service.bundler().calls == 2);
    std::printf("  landmarks reset by the winsorizer: %d\n", service.last_winsorized());
  }

  std::printf("%s (%d failures)\n", g_failures ? "FAILED" : "OK", g_failures);
  return g_failures ? 1 : 0;
}
