// Synthetic sliding-window scenes for the SBA benchmark configurations.
//
// A pinhole camera (fx, fy, cx, cy, width x height) on key frames spaced along +x, each with a small random
// rotation; points uniform in a box in front of the cameras. Observations are pixel projections with Gaussian
// noise and a fraction of uniform outliers, converted to the normalized image coordinates the solver works in
// (the reference's SBA has no intrinsics: libs/sba/schur_complement_bundler_cpu.cpp:98). The information
// matrix is diag(fx^2, fy^2) / huber_px^2, the pixel-space information mapped to normalized coordinates the
// way libs/camera/observation.cpp:57-67 does, so a squared Mahalanobis residual is in px^2.
//
// Storage follows the reference's conventions: the first key frame (the gauge) is stored last and is the one
// fixed key frame (bundle_adjustment_problem.h:52-56, pipelines/service_sba.h:118-134); observations are
// ordered by key frame, then by point id (bundle_adjustment_problem.h:31-33).
//
// The random stream is this file's own (splitmix64-seeded xoshiro256**), so a seed names the same scene on
// every machine.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cuvslam_sba.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& w : s) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      w = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int uniform_int(int lo, int hi) {  // inclusive
    return lo + static_cast<int>(uniform() * (hi - lo + 1));
  }
  bool have_spare = false;
  double spare = 0;
  double normal() {  // Box-Muller
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    double u1 = uniform();
    while (u1 <= 1e-300) u1 = uniform();
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2.0 * M_PI * u2);
    have_spare = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
};

struct Pose {
  double R[3][3];
  double t[3];
};

// Rodrigues rotation of the axis-angle vector w.
void rot_exp(const double w[3], double R[3][3]) {
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  if (th < 1e-15) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[i][j] = (i == j);
    return;
  }
  const double n[3] = {w[0] / th, w[1] / th, w[2] / th};
  const double c = std::cos(th), s = std::sin(th);
  const double K[3][3] = {{0, -n[2], n[1]}, {n[2], 0, -n[0]}, {-n[1], n[0], 0}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = c * (i == j) + s * K[i][j] + (1 - c) * n[i] * n[j];
}

void matmul3(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

void apply(const Pose& T, const double p[3], double q[3]) {
  for (int i = 0; i < 3; ++i) q[i] = T.R[i][0] * p[0] + T.R[i][1] * p[1] + T.R[i][2] * p[2] + T.t[i];
}

void store_pose(const Pose& T, float* m) {  // column-major 4x4 (Eigen Isometry3f layout)
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[c * 4 + r] = static_cast<float>(T.R[r][c]);
  m[3] = m[7] = m[11] = 0.f;
  for (int r = 0; r < 3; ++r) m[12 + r] = static_cast<float>(T.t[r]);
  m[15] = 1.f;
}

}  // namespace

struct sba_scene {
  std::vector<float> points, gt_points;
  std::vector<float> obs_xy, obs_info;
  std::vector<int32_t> point_ids, pose_ids;
  std::vector<int8_t> camera_ids;
  std::vector<float> poses, gt_poses;
  std::vector<float> cams;
  int32_t num_fixed_kf = 1;
};

extern "C" void sba_scene_default_config(sba_scene_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->num_keyframes = 10;
  c->num_points = 2000;
  c->sample_observations = 0;
  c->num_cameras = 1;
  c->num_fixed_keyframes = 1;
  c->fx = c->fy = 450.0;
  c->cx = 376.0;
  c->cy = 240.0;
  c->width = 752;
  c->height = 480;
  c->keyframe_spacing = 0.1;
  c->keyframe_rotation_deg = 1.0;
  c->pixel_noise = 0.5;
  c->outlier_rate = 0.05;
  c->pose_sigma_m = 0.02;
  c->pose_sigma_deg = 0.5;
  c->point_sigma_m = 0.05;
  c->huber_px = 1.0;
  c->xy_extent = 5.0;
  c->z_min = 4.0;
  c->z_max = 20.0;
  c->seed = 1;
}

extern "C" int sba_scene_generate(const sba_scene_config* cfg, sba_scene** out) {
  if (!cfg || !out || cfg->num_keyframes < 1 || cfg->num_points < 0 || cfg->num_cameras != 1 || cfg->num_fixed_keyframes < 1) return SBA_E_INVALID;
  sba_scene* sc = new (std::nothrow) sba_scene;
  if (!sc) return SBA_E_INVALID;
  Rng rng(cfg->seed);
  const double deg = M_PI / 180.0;
  const int nk = cfg->num_keyframes;

  // Ground-truth key frames in generation order k = 0..nk-1; rig_from_world = (world_from_rig)^-1.
  std::vector<Pose> gt(nk);
  for (int k = 0; k < nk; ++k) {
    double w[3] = {rng.normal() * cfg->keyframe_rotation_deg * deg, rng.normal() * cfg->keyframe_rotation_deg * deg,
                   rng.normal() * cfg->keyframe_rotation_deg * deg};
    double Rwr[3][3];
    rot_exp(w, Rwr);
    const double c[3] = {cfg->keyframe_spacing * k, 0.0, 0.0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gt[k].R[i][j] = Rwr[j][i];
    for (int i = 0; i < 3; ++i) gt[k].t[i] = -(gt[k].R[i][0] * c[0] + gt[k].R[i][1] * c[1] + gt[k].R[i][2] * c[2]);
  }
  // Stored order: the first nf generated key frames (the fixed gauge frames) go last.
  const int nf = std::max(1, std::min(cfg->num_fixed_keyframes, nk));
  auto stored = [nk, nf](int k) { return k < nf ? nk - nf + k : k - nf; };

  auto project = [&](const Pose& T, const double p[3], double& u, double& v) -> bool {
    double q[3];
    apply(T, p, q);
    if (!(q[2] > 0.1)) return false;
    u = cfg->fx * q[0] / q[2] + cfg->cx;
    v = cfg->fy * q[1] / q[2] + cfg->cy;
    return u >= 0.0 && u < cfg->width && v >= 0.0 && v < cfg->height;
  };

  // Points and, per point, the (generation-order) key frames that observe it.
  const int np = cfg->num_points;
  std::vector<double> P(3 * static_cast<size_t>(np));
  std::vector<std::vector<int>> seen(np);
  std::vector<int> vis;
  for (int i = 0; i < np; ++i) {
    double p[3];
    for (int attempt = 0;; ++attempt) {
      p[0] = rng.uniform(-cfg->xy_extent, cfg->xy_extent);
      p[1] = rng.uniform(-cfg->xy_extent, cfg->xy_extent);
      p[2] = rng.uniform(cfg->z_min, cfg->z_max);
      vis.clear();
      for (int k = 0; k < nk; ++k) {
        double u, v;
        if (project(gt[k], p, u, v)) vis.push_back(k);
      }
      // Sampled windows keep only points that two key frames can see; bounded retries keep this finite.
      if (!cfg->sample_observations || vis.size() >= 2 || attempt >= 64) break;
    }
    if (cfg->sample_observations && vis.size() >= 2) {
      const int m = std::min<int>(rng.uniform_int(2, nk), static_cast<int>(vis.size()));
      for (int a = 0; a < m; ++a) {  // partial Fisher-Yates: m key frames without replacement
        const int b = a + static_cast<int>(rng.uniform() * (vis.size() - a));
        std::swap(vis[a], vis[b]);
      }
      vis.resize(m);
      std::sort(vis.begin(), vis.end());
    } else if (cfg->sample_observations) {
      vis.clear();
    }
    seen[i] = vis;
    for (int d = 0; d < 3; ++d) P[3 * i + d] = p[d];
  }

  // Observations ordered by stored key frame, then by point id.
  std::vector<std::vector<int>> kf_points(nk);  // indexed by stored key frame
  for (int i = 0; i < np; ++i)
    for (int k : seen[i]) kf_points[stored(k)].push_back(i);
  std::vector<int> gen_of_stored(nk);
  for (int k = 0; k < nk; ++k) gen_of_stored[stored(k)] = k;

  const float info_x = static_cast<float>(cfg->fx * cfg->fx / (cfg->huber_px * cfg->huber_px));
  const float info_y = static_cast<float>(cfg->fy * cfg->fy / (cfg->huber_px * cfg->huber_px));
  for (int s = 0; s < nk; ++s) {
    const Pose& T = gt[gen_of_stored[s]];
    for (int i : kf_points[s]) {
      double u, v;
      project(T, &P[3 * i], u, v);
      if (rng.uniform() < cfg->outlier_rate) {
        u = rng.uniform(0.0, cfg->width);
        v = rng.uniform(0.0, cfg->height);
      } else {
        u += rng.normal() * cfg->pixel_noise;
        v += rng.normal() * cfg->pixel_noise;
      }
      sc->obs_xy.push_back(static_cast<float>((u - cfg->cx) / cfg->fx));
      sc->obs_xy.push_back(static_cast<float>((v - cfg->cy) / cfg->fy));
      const float info[4] = {info_x, 0.f, 0.f, info_y};
      sc->obs_info.insert(sc->obs_info.end(), info, info + 4);
      sc->point_ids.push_back(i);
      sc->pose_ids.push_back(s);
      sc->camera_ids.push_back(0);
    }
  }

  // Ground truth and the perturbed initial state (free key frames and all points move).
  sc->gt_points.resize(3 * static_cast<size_t>(np));
  sc->points.resize(3 * static_cast<size_t>(np));
  for (int i = 0; i < np; ++i)
    for (int d = 0; d < 3; ++d) {
      sc->gt_points[3 * i + d] = static_cast<float>(P[3 * i + d]);
      sc->points[3 * i + d] = static_cast<float>(P[3 * i + d] + rng.normal() * cfg->point_sigma_m);
    }
  sc->gt_poses.resize(16 * static_cast<size_t>(nk));
  sc->poses.resize(16 * static_cast<size_t>(nk));
  for (int s = 0; s < nk; ++s) {
    const Pose& T = gt[gen_of_stored[s]];
    store_pose(T, &sc->gt_poses[16 * s]);
    Pose Tp = T;
    if (s < nk - nf) {  // free key frame: left-multiply by a random rigid motion
      double w[3] = {rng.normal() * cfg->pose_sigma_deg * deg, rng.normal() * cfg->pose_sigma_deg * deg,
                     rng.normal() * cfg->pose_sigma_deg * deg};
      double dR[3][3];
      rot_exp(w, dR);
      const double dt[3] = {rng.normal() * cfg->pose_sigma_m, rng.normal() * cfg->pose_sigma_m,
                            rng.normal() * cfg->pose_sigma_m};
      matmul3(dR, T.R, Tp.R);
      for (int i = 0; i < 3; ++i)
        Tp.t[i] = dR[i][0] * T.t[0] + dR[i][1] * T.t[1] + dR[i][2] * T.t[2] + dt[i];
    }
    store_pose(Tp, &sc->poses[16 * s]);
  }
  sc->cams.assign(16, 0.f);
  sc->cams[0] = sc->cams[5] = sc->cams[10] = sc->cams[15] = 1.f;
  sc->num_fixed_kf = nf;
  *out = sc;
  return SBA_OK;
}

extern "C" int sba_scene_view(const sba_scene* s, sba_problem_view* v) {
  if (!s || !v) return SBA_E_INVALID;
  v->num_points = static_cast<int32_t>(s->points.size() / 3);
  v->num_fixed_points = 0;
  v->num_observations = static_cast<int32_t>(s->point_ids.size());
  v->num_poses = static_cast<int32_t>(s->poses.size() / 16);
  v->num_fixed_key_frames = s->num_fixed_kf;
  v->num_cameras = static_cast<int32_t>(s->cams.size() / 16);
  v->points = s->points.data();
  v->observation_xys = s->obs_xy.data();
  v->observation_infos = s->obs_info.data();
  v->point_ids = s->point_ids.data();
  v->pose_ids = s->pose_ids.data();
  v->camera_ids = s->camera_ids.data();
  v->rig_from_world = s->poses.data();
  v->camera_from_rig = s->cams.data();
  v->max_iterations = 10;
  v->robustifier_scale = 1.f;
  return SBA_OK;
}

extern "C" int sba_scene_ground_truth(const sba_scene* s, const float** points, const float** poses) {
  if (!s) return SBA_E_INVALID;
  if (points) *points = s->gt_points.data();
  if (poses) *poses = s->gt_poses.data();
  return SBA_OK;
}

extern "C" void sba_scene_free(sba_scene* s) { delete s; }

// ---------------------------------------------------------------------------------------------------------
// The large multi-camera map: a stereo rig on a closed loop inside a corridor.
// ---------------------------------------------------------------------------------------------------------

extern "C" void sba_loop_default_config(sba_loop_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->num_keyframes = 200;
  c->num_points = 200000;
  c->num_stereo_pairs = 4;
  c->num_fixed_keyframes = 1;
  c->loop_length_m = 20.0;
  c->stereo_baseline_m = 0.12;
  c->corridor_half_m = 1.5;
  c->floor_y_m = 1.2;     // y points down in the rig and camera frames
  c->ceiling_y_m = -1.5;
  c->clutter_fraction = 0.3;
  c->max_range_m = 1.6;
  c->fx = c->fy = 450.0;
  c->cx = 376.0;
  c->cy = 240.0;
  c->width = 752;
  c->height = 480;
  c->keyframe_rotation_deg = 1.0;
  c->pixel_noise = 0.5;
  c->outlier_rate = 0.03;
  c->pose_sigma_m = 0.02;
  c->pose_sigma_deg = 0.5;
  c->point_sigma_m = 0.05;
  c->huber_px = 1.0;
  c->seed = 5;
}

extern "C" int sba_scene_generate_loop(const sba_loop_config* cfg, sba_scene** out) {
  if (!cfg || !out || cfg->num_keyframes < 2 || cfg->num_points < 0 || cfg->num_stereo_pairs < 1 ||
      2 * cfg->num_stereo_pairs > 32 || cfg->num_fixed_keyframes < 1 || cfg->max_range_m <= 0.1)
    return SBA_E_INVALID;
  sba_scene* sc = new (std::nothrow) sba_scene;
  if (!sc) return SBA_E_INVALID;
  Rng rng(cfg->seed);
  const double deg = M_PI / 180.0;
  const int nk = cfg->num_keyframes;
  const double r = cfg->loop_length_m / (2.0 * M_PI);
  // key frames: rig at angle a on the loop, z (forward) along the tangent, y down, x = y cross z
  std::vector<Pose> gt(nk);
  for (int k = 0; k < nk; ++k) {
    const double a = 2.0 * M_PI * k / nk;
    const double cpos[3] = {r * std::cos(a), 0.0, r * std::sin(a)};
    const double zf[3] = {-std::sin(a), 0.0, std::cos(a)};
    const double yd[3] = {0.0, 1.0, 0.0};
    const double xr[3] = {yd[1] * zf[2] - yd[2] * zf[1], yd[2] * zf[0] - yd[0] * zf[2], yd[0] * zf[1] - yd[1] * zf[0]};
    double Rwr0[3][3];  // columns: rig axes in the world
    for (int i = 0; i < 3; ++i) {
      Rwr0[i][0] = xr[i];
      Rwr0[i][1] = yd[i];
      Rwr0[i][2] = zf[i];
    }
    double w[3] = {rng.normal() * cfg->keyframe_rotation_deg * deg, rng.normal() * cfg->keyframe_rotation_deg * deg,
                   rng.normal() * cfg->keyframe_rotation_deg * deg};
    double J[3][3], Rwr[3][3];
    rot_exp(w, J);
    matmul3(Rwr0, J, Rwr);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gt[k].R[i][j] = Rwr[j][i];
    for (int i = 0; i < 3; ++i)
      gt[k].t[i] = -(gt[k].R[i][0] * cpos[0] + gt[k].R[i][1] * cpos[1] + gt[k].R[i][2] * cpos[2]);
  }
  // rig: stereo pair q looks along yaw 2 pi q / n (about the rig's y axis); its cameras sit +-baseline/2 along
  // the camera's own x axis
  const int ncam = 2 * cfg->num_stereo_pairs;
  std::vector<Pose> cam(ncam);
  for (int c = 0; c < ncam; ++c) {
    const double yaw = 2.0 * M_PI * (c / 2) / cfg->num_stereo_pairs;
    const double side = (c % 2 ? 0.5 : -0.5) * cfg->stereo_baseline_m;
    // rotation rig -> camera: R_cr = Ry(yaw)^T
    const double cy_ = std::cos(yaw), sy = std::sin(yaw);
    const double Rrc[3][3] = {{cy_, 0, sy}, {0, 1, 0}, {-sy, 0, cy_}};  // camera axes in the rig frame
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) cam[c].R[i][j] = Rrc[j][i];
    const double center[3] = {Rrc[0][0] * side, Rrc[1][0] * side, Rrc[2][0] * side};
    for (int i = 0; i < 3; ++i)
      cam[c].t[i] = -(cam[c].R[i][0] * center[0] + cam[c].R[i][1] * center[1] + cam[c].R[i][2] * center[2]);
  }
  const int nf = std::max(1, std::min(cfg->num_fixed_keyframes, nk));
  auto stored = [nk, nf](int k) { return k < nf ? nk - nf + k : k - nf; };
  auto sees = [&](int k, int c, const double p[3], double& u, double& v) -> bool {
    double q[3], s[3];
    apply(gt[k], p, q);
    apply(cam[c], q, s);
    if (!(s[2] > 0.1)) return false;
    if (s[0] * s[0] + s[1] * s[1] + s[2] * s[2] > cfg->max_range_m * cfg->max_range_m) return false;
    u = cfg->fx * s[0] / s[2] + cfg->cx;
    v = cfg->fy * s[1] / s[2] + cfg->cy;
    return u >= 0.0 && u < cfg->width && v >= 0.0 && v < cfg->height;
  };
  const double band = cfg->corridor_half_m;
  auto sample_point = [&](double p[3], double& theta) {
    theta = rng.uniform(0.0, 2.0 * M_PI);
    double rad, y;
    if (rng.uniform() >= cfg->clutter_fraction) {
      const int face = rng.uniform_int(0, 3);
      rad = face == 0 ? r - band : face == 1 ? r + band : rng.uniform(r - band, r + band);
      y = face == 2 ? cfg->floor_y_m : face == 3 ? cfg->ceiling_y_m : rng.uniform(cfg->ceiling_y_m, cfg->floor_y_m);
    } else {
      double off = rng.uniform(0.4, band);  // clutter keeps 0.4 m off the path
      rad = rng.uniform() < 0.5 ? r - off : r + off;
      y = rng.uniform(cfg->ceiling_y_m, cfg->floor_y_m);
    }
    p[0] = rad * std::cos(theta);
    p[1] = y;
    p[2] = rad * std::sin(theta);
  };
  // points with at least two key frames that see them, in order of angle
  const int np = cfg->num_points;
  struct Pt {
    double theta, p[3];
    std::vector<std::pair<int, int>> views;  // (generation key frame, camera)
  };
  std::vector<Pt> pts(np);
  std::vector<int> kfs;
  for (int i = 0; i < np; ++i) {
    Pt& q = pts[i];
    for (int attempt = 0;; ++attempt) {
      sample_point(q.p, q.theta);
      q.views.clear();
      kfs.clear();
      // key frames near the point's angle only: the range bounds visibility to a short arc
      const int k0 = static_cast<int>(std::floor(q.theta / (2.0 * M_PI) * nk));
      const int span = static_cast<int>(std::ceil((cfg->max_range_m + band) / cfg->loop_length_m * nk)) + 2;
      for (int dk = -span; dk <= span; ++dk) {
        const int k = ((k0 + dk) % nk + nk) % nk;
        bool any = false;
        for (int c = 0; c < ncam; ++c) {
          double u, v;
          if (sees(k, c, q.p, u, v)) {
            q.views.emplace_back(k, c);
            any = true;
          }
        }
        if (any) kfs.push_back(k);
      }
      if (kfs.size() >= 2 || attempt >= 256) break;
    }
    if (kfs.size() < 2) {
      q.views.clear();
      continue;
    }
    // keep U{2..n} of the n key frames
    const int m = rng.uniform_int(2, static_cast<int>(kfs.size()));
    for (int a = 0; a < m; ++a) {
      const int b = a + static_cast<int>(rng.uniform() * (kfs.size() - a));
      std::swap(kfs[a], kfs[b]);
    }
    kfs.resize(m);
    std::vector<std::pair<int, int>> keep;
    for (auto& vw : q.views)
      if (std::find(kfs.begin(), kfs.end(), vw.first) != kfs.end()) keep.push_back(vw);
    q.views.swap(keep);
  }
  std::sort(pts.begin(), pts.end(), [](const Pt& a, const Pt& b) { return a.theta < b.theta; });
  // observations ordered by stored key frame, then point id, then camera
  std::vector<std::vector<std::pair<int, int>>> kf_obs(nk);  // stored key frame -> (point, camera)
  for (int i = 0; i < np; ++i)
    for (auto& vw : pts[i].views) kf_obs[stored(vw.first)].emplace_back(i, vw.second);
  std::vector<int> gen_of_stored(nk);
  for (int k = 0; k < nk; ++k) gen_of_stored[stored(k)] = k;
  const float info_x = static_cast<float>(cfg->fx * cfg->fx / (cfg->huber_px * cfg->huber_px));
  const float info_y = static_cast<float>(cfg->fy * cfg->fy / (cfg->huber_px * cfg->huber_px));
  for (int s = 0; s < nk; ++s) {
    std::sort(kf_obs[s].begin(), kf_obs[s].end());
    for (auto& pc : kf_obs[s]) {
      const int i = pc.first, c = pc.second;
      double u = 0, v = 0;
      sees(gen_of_stored[s], c, pts[i].p, u, v);
      if (rng.uniform() < cfg->outlier_rate) {
        u = rng.uniform(0.0, cfg->width);
        v = rng.uniform(0.0, cfg->height);
      } else {
        u += rng.normal() * cfg->pixel_noise;
        v += rng.normal() * cfg->pixel_noise;
      }
      sc->obs_xy.push_back(static_cast<float>((u - cfg->cx) / cfg->fx));
      sc->obs_xy.push_back(static_cast<float>((v - cfg->cy) / cfg->fy));
      const float info[4] = {info_x, 0.f, 0.f, info_y};
      sc->obs_info.insert(sc->obs_info.end(), info, info + 4);
      sc->point_ids.push_back(i);
      sc->pose_ids.push_back(s);
      sc->camera_ids.push_back(static_cast<int8_t>(c));
    }
  }
  sc->gt_points.resize(3 * static_cast<size_t>(np));
  sc->points.resize(3 * static_cast<size_t>(np));
  for (int i = 0; i < np; ++i)
    for (int d = 0; d < 3; ++d) {
      sc->gt_points[3 * i + d] = static_cast<float>(pts[i].p[d]);
      sc->points[3 * i + d] = static_cast<float>(pts[i].p[d] + rng.normal() * cfg->point_sigma_m);
    }
  sc->gt_poses.resize(16 * static_cast<size_t>(nk));
  sc->poses.resize(16 * static_cast<size_t>(nk));
  for (int s = 0; s < nk; ++s) {
    const Pose& T = gt[gen_of_stored[s]];
    store_pose(T, &sc->gt_poses[16 * s]);
    Pose Tp = T;
    if (s < nk - nf) {
      double w[3] = {rng.normal() * cfg->pose_sigma_deg * deg, rng.normal() * cfg->pose_sigma_deg * deg,
                     rng.normal() * cfg->pose_sigma_deg * deg};
      double dR[3][3];
      rot_exp(w, dR);
      const double dt[3] = {rng.normal() * cfg->pose_sigma_m, rng.normal() * cfg->pose_sigma_m,
                            rng.normal() * cfg->pose_sigma_m};
      matmul3(dR, T.R, Tp.R);
      for (int i = 0; i < 3; ++i) Tp.t[i] = dR[i][0] * T.t[0] + dR[i][1] * T.t[1] + dR[i][2] * T.t[2] + dt[i];
    }
    store_pose(Tp, &sc->poses[16 * s]);
  }
  sc->cams.resize(16 * static_cast<size_t>(ncam));
  for (int c = 0; c < ncam; ++c) store_pose(cam[c], &sc->cams[16 * c]);
  sc->num_fixed_kf = nf;
  *out = sc;
  return SBA_OK;
}
