// A caller of the reference's API, compiled against the drop-in headers: fills a
// cuvslam::sba::BundleAdjustmentProblem from a synthetic scene and calls SchurComplementBundlerGpu::solve,
// exactly as pipelines/service_sba.h:168 does. Prints the solver statistics as one JSON line.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cuvslam_sba.h"
#include "sba/schur_complement_bundler_gpu.h"

int main(int argc, char** argv) {
  const int seed = argc > 1 ? std::atoi(argv[1]) : 1;
  const int loss = argc > 2 ? std::atoi(argv[2]) : 1;
  sba_scene_config cfg;
  sba_scene_default_config(&cfg);
  cfg.seed = static_cast<uint64_t>(seed);
  cfg.num_points = 400;
  cfg.num_keyframes = 6;
  cfg.num_fixed_keyframes = 2;
  sba_scene* sc = nullptr;
  if (sba_scene_generate(&cfg, &sc) != SBA_OK) return 2;
  sba_problem_view v;
  sba_scene_view(sc, &v);

  cuvslam::sba::BundleAdjustmentProblem p;
  p.points.resize(v.num_points);
  std::memcpy(static_cast<void*>(p.points.data()), v.points, sizeof(float) * 3 * v.num_points);
  p.observation_xys.resize(v.num_observations);
  std::memcpy(static_cast<void*>(p.observation_xys.data()), v.observation_xys, sizeof(float) * 2 * v.num_observations);
  p.observation_infos.resize(v.num_observations);
  std::memcpy(static_cast<void*>(p.observation_infos.data()), v.observation_infos, sizeof(float) * 4 * v.num_observations);
  p.point_ids.assign(v.point_ids, v.point_ids + v.num_observations);
  p.pose_ids.assign(v.pose_ids, v.pose_ids + v.num_observations);
  p.camera_ids.assign(v.camera_ids, v.camera_ids + v.num_observations);
  p.rig_from_world.resize(v.num_poses);
  std::memcpy(static_cast<void*>(p.rig_from_world.data()), v.rig_from_world, sizeof(float) * 16 * v.num_poses);
  p.num_fixed_key_frames = v.num_fixed_key_frames;
  p.rig.num_cameras = v.num_cameras;
  std::memcpy(static_cast<void*>(p.rig.camera_from_rig[0].data()), v.camera_from_rig, sizeof(float) * 16);
  p.max_iterations = 10;
  p.robustifier_scale = 1.f;
  sba_scene_free(sc);

  cuvslam::sba::SchurComplementBundlerGpu bundler(loss);
  const bool ok = bundler.solve(p);
  std::printf("{\"ok\": %s, \"iterations\": %d, \"initial_cost\": %.9g, \"last_cost\": %.9g, \"p0\": [%.9g, %.9g, %.9g]}\n",
              ok ? "true" : "false", p.iterations, p.initial_cost, p.last_cost, p.points[0].x(), p.points[0].y(),
              p.points[0].z());
  return ok ? 0 : 1;
}
