// CPU run of the mono SBA solver body (cuvslam_amd/csrc/sba_mono_body.h) with one "thread": the arithmetic the
// GPU workgroup performs, sequentially, for comparison with the oracle before the kernel runs on a GPU. Reads one
// window (little endian: int32 F T O; int32 constrained[F]; float poses[F][16] (world_from_camera, column-major),
// points[T][3]; int32 obs_frame[O], obs_track[O]; float obs_xy[O][2]; int32 max_iterations; float min_residual,
// min_residual_speed) and prints the result as one JSON line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "../../cuvslam_amd/csrc/sba_mono_body.h"
#include "../../cuvslam_amd/csrc/sba_mono_pack.h"

using namespace sba_mono;

template <class T>
static bool get(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return std::fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int32_t> hdr, con, of, ot, mi;
  std::vector<float> poses, pts, xy, mr;
  bool ok = get(f, hdr, 3);
  const int F = ok ? hdr[0] : 0, T = ok ? hdr[1] : 0, O = ok ? hdr[2] : 0;
  ok = ok && get(f, con, F) && get(f, poses, 16 * size_t(F)) && get(f, pts, 3 * size_t(T)) && get(f, of, O) &&
       get(f, ot, O) && get(f, xy, 2 * size_t(O)) && get(f, mi, 1) && get(f, mr, 2);
  std::fclose(f);
  if (!ok) return 3;
  sba_mono_window v{F, T, O, con.data(), poses.data(), pts.data(), of.data(), ot.data(), xy.data(), mi[0], mr[0],
                    mr[1]};
  MonoPack k;
  if (!pack_window(v, k)) return 4;
  const int n = 6 * k.nFree;
  std::vector<double> fa(6 * F), faD(6 * F), fda(6 * F), ferrA(6 * F), feps(6 * F), fUo(36 * F), fUa(36 * F),
      fM(12 * F), fRn(9 * F), tb(3 * T), tbD(3 * T), tdb(3 * T), terrB(3 * T), tV(9 * T), tinvVa(9 * T), tinfo(4 * T),
      oerr(2 * O), ow(O), oA(12 * O), oB(6 * O), oW(18 * O), oY(18 * O), S(n * n + 1), Sb(n + 1), Q(n * n + 1),
      E(n * n + 1), E2(n * n + 1), ev(n + 1), x(n + 1), xt(n + 1), bd(n + 1), dv(n + 1), rot(2 * (n + 2)), part(kPart * std::max({kMonoThreads, F, k.nFree * k.nFree})), prep(36), cov(36), sc(8);
  std::vector<int32_t> tskip(T);
  MonoWin w{};
  w.nF = F; w.nT = T; w.nO = O; w.nFree = k.nFree; w.max_iterations = mi[0];
  w.min_residual = mr[0]; w.min_residual_speed = mr[1];
  w.constrained = k.constrained.data(); w.free_idx = k.free_idx.data(); w.pose_in = poses.data();
  w.fa = fa.data(); w.faD = faD.data(); w.fda = fda.data(); w.ferrA = ferrA.data(); w.feps = feps.data();
  w.fUo = fUo.data(); w.fUa = fUa.data(); w.fM = fM.data(); w.fRn = fRn.data();
  w.frame_obs = k.frame_obs.data(); w.fobs = k.fobs.data(); w.pts_in = pts.data();
  w.track_obs = k.track_obs.data(); w.tobs = k.tobs.data(); w.tskip = tskip.data(); w.tb = tb.data();
  w.tbD = tbD.data(); w.tdb = tdb.data(); w.terrB = terrB.data(); w.tV = tV.data(); w.tinvVa = tinvVa.data();
  w.tinfo = tinfo.data(); w.oframe = k.oframe.data(); w.otrack = k.otrack.data(); w.oxy = xy.data();
  w.oerr = oerr.data(); w.ow = ow.data(); w.oA = oA.data(); w.oB = oB.data(); w.oW = oW.data(); w.oY = oY.data();
  w.blk = k.blk.data(); w.bpair = reinterpret_cast<const int2*>(k.bpair.data()); w.S = S.data(); w.Sb = Sb.data();
  w.Q = Q.data(); w.E = E.data(); w.E2 = E2.data(); w.ev = ev.data(); w.x = x.data(); w.xt = xt.data(); w.bd = bd.data();
  w.dv = dv.data(); w.rot = rot.data(); w.part = part.data(); w.prep = prep.data(); w.cov = cov.data(); w.out_scalars = sc.data();
  static MonoShared sh;
  MonoSolve s{w, sh, 0, 1};
  s.run();
  std::vector<float> out_poses(16 * size_t(F)), out_pts(3 * size_t(T));
  s.export_state(out_poses.data(), out_pts.data());
  std::printf("{\"residual\": %.17g, \"residual_initial\": %.17g, \"lambda\": %.17g, \"iterations\": %d, "
              "\"failed\": %d, \"feasible\": %d, \"poses\": [", sc[0], sc[1], sc[2], int(sc[3]), int(sc[4]),
              int(sc[5]));
  for (int i = 0; i < 16 * F; ++i) std::printf("%s%.9g", i ? ", " : "", out_poses[i]);
  std::printf("], \"points\": [");
  for (int i = 0; i < 3 * T; ++i) std::printf("%s%.9g", i ? ", " : "", out_pts[i]);
  std::printf("], \"cov\": [");
  for (int i = 0; i < 36; ++i) std::printf("%s%.17g", i ? ", " : "", cov[i]);
  std::printf("]}\n");
  return 0;
}
