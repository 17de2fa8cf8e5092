// SchurComplementBundlerGpu on the MI355X SBA (the class of libs/sba/schur_complement_bundler_gpu.h:29-38).
// The problem's vectors are handed to the C ABI as the plain arrays they already are.
#include "sba/schur_complement_bundler_gpu.h"

#include "cuvslam_sba.h"

namespace cuvslam::sba {

static_assert(sizeof(Vector3T) == 12, "Vector3T must be three packed floats");
static_assert(sizeof(Vector2T) == 8, "Vector2T must be two packed floats");
static_assert(sizeof(Matrix2T) == 16, "Matrix2T must be four floats");
static_assert(sizeof(Isometry3T) == 64, "Isometry3T must be a 4x4 float matrix");

class SchurComplementBundlerGpu::Impl {
public:
  Impl(int loss, int device) {
    sba_options o;
    sba_default_options(&o);
    o.loss = loss;
    o.device = device;
    if (sba_create(&o, &h) != SBA_OK) h = nullptr;
  }
  ~Impl() { sba_destroy(h); }
  sba_handle* h = nullptr;
};

SchurComplementBundlerGpu::SchurComplementBundlerGpu() : impl(new Impl(SBA_LOSS_STUDENT_T, 0)) {}
SchurComplementBundlerGpu::SchurComplementBundlerGpu(int loss, int device) : impl(new Impl(loss, device)) {}
SchurComplementBundlerGpu::~SchurComplementBundlerGpu() = default;

bool SchurComplementBundlerGpu::solve(BundleAdjustmentProblem& p) {
  if (!impl || !impl->h) return false;
  sba_problem_view v{};
  v.num_points = static_cast<int32_t>(p.points.size());
  v.num_fixed_points = p.num_fixed_points;
  v.num_observations = static_cast<int32_t>(p.observation_xys.size());
  v.num_poses = static_cast<int32_t>(p.rig_from_world.size());
  v.num_fixed_key_frames = p.num_fixed_key_frames;
  v.num_cameras = p.rig.num_cameras;
  if (p.observation_infos.size() != p.observation_xys.size() || p.point_ids.size() != p.observation_xys.size() ||
      p.pose_ids.size() != p.observation_xys.size() || p.camera_ids.size() != p.observation_xys.size())
    return false;
  v.points = reinterpret_cast<const float*>(p.points.data());
  v.observation_xys = reinterpret_cast<const float*>(p.observation_xys.data());
  v.observation_infos = reinterpret_cast<const float*>(p.observation_infos.data());
  v.point_ids = p.point_ids.data();
  v.pose_ids = p.pose_ids.data();
  v.camera_ids = p.camera_ids.data();
  v.rig_from_world = reinterpret_cast<const float*>(p.rig_from_world.data());
  v.camera_from_rig = reinterpret_cast<const float*>(p.rig.camera_from_rig);
  v.max_iterations = p.max_iterations;
  v.robustifier_scale = p.robustifier_scale;
  sba_result r{};
  r.points = reinterpret_cast<float*>(p.points.data());
  r.rig_from_world = reinterpret_cast<float*>(p.rig_from_world.data());
  const int rc = sba_solve(impl->h, &v, &r);
  if (rc == SBA_E_EMPTY || rc == SBA_E_DEVICE || rc == SBA_E_CAPACITY || rc == SBA_E_INVALID) return false;
  p.iterations = r.iterations;
  p.initial_cost = r.initial_cost;
  p.last_cost = r.last_cost;
  return rc == SBA_OK;
}

}  // namespace cuvslam::sba
