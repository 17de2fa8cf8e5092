// TEST INFRASTRUCTURE -- runs the CPU oracle's entry points on generated windows so that the oracle can be built and
// checked under AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile: asan_driver). Exits non-zero on a
// sanitizer report (they abort) or on a failed status.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sba_oracle.cpp"  // the oracle's own translation unit, for its types
#include "../include/cuvslam_sba.h"

int main() {
  int bad = 0;
  for (int w = 0; w < 3; ++w) {
    sba_scene_config cfg;
    sba_scene_default_config(&cfg);
    cfg.seed = 100 + w;
    cfg.num_keyframes = 4 + 3 * w;
    cfg.num_points = 150 + 100 * w;
    cfg.num_fixed_keyframes = 2;
    cfg.sample_observations = w == 2;
    sba_scene* sc = nullptr;
    if (sba_scene_generate(&cfg, &sc) != 0) return 2;
    sba_problem_view v;
    sba_scene_view(sc, &v);
    v.max_iterations = 6;
    const OrcProblem* P = reinterpret_cast<const OrcProblem*>(&v);
    const int np = v.num_points, nc = v.num_poses, no = v.num_observations;
    const int nx = 6 * (nc - v.num_fixed_key_frames) + 3 * (np - v.num_fixed_points);
    std::vector<double> pts(3 * np), poses(16 * nc), tl(16), tc(16), tr(16);
    std::vector<int32_t> tf(16);
    for (int loss = 0; loss < 2; ++loss) {
      OrcOptions o{};
      o.loss = loss;
      o.reorthonormalize = 1;
      OrcResult r{};
      r.points = pts.data();
      r.poses = poses.data();
      r.trace_capacity = 16;
      r.trace_lambda = tl.data();
      r.trace_cost = tc.data();
      r.trace_rho = tr.data();
      r.trace_flags = tf.data();
      bad += orc_solve_f64(P, &o, &r) != 0 || !std::isfinite(r.last_cost);
      o.rcs_solver = 1;
      o.dense_schur = 1;
      bad += orc_solve_f32(P, &o, &r) != 0;
      o.rcs_solver = 0;
      o.dense_schur = 0;
      o.round_reduced_f32 = 2;
      o.flip_iteration = 2;
      bad += orc_solve_f64(P, &o, &r) != 0;
      o.round_reduced_f32 = 0;
      o.flip_iteration = 0;
      const int n6 = 6 * (nc - v.num_fixed_key_frames);
      std::vector<double> S(static_cast<size_t>(n6) * n6), e(n6), dc(n6), dp(3 * (np - v.num_fixed_points)), sc9(9);
      orc_step_f64(P, &o, 1e-2, S.data(), e.data(), dc.data(), dp.data(), sc9.data());
      std::vector<double> g(nx), x(nx, 1e-3), y(nx), a(nx), dxy(2 * no), dinfo(3 * no);
      orc_grad_f64(P, &o, nullptr, nullptr, nullptr, nullptr, g.data());
      orc_hvp_f64(P, &o, nullptr, nullptr, x.data(), y.data());
      orc_implicit_grad_f64(P, &o, nullptr, nullptr, nullptr, nullptr, g.data(), a.data(), dxy.data(), dinfo.data(), 20,
                            1e-10);
      std::vector<uint8_t> reset(np);
      orc_winsorize_f64(P, nc - 1, nullptr, nullptr, reset.data());
    }
    sba_scene_free(sc);
  }
  std::printf("%s\n", bad ? "FAILED" : "OK");
  return bad ? 1 : 0;
}
