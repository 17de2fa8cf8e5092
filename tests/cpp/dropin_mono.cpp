// A caller of the reference's MonoSBASolver API, compiled against the drop-in header: reads a window in the format
// of mono_host_check.cpp, adds its tracks (addTrack) and frames with their observations (addFrame, track ids
// starting at 1 as valid reference ids do), solves on the GPU and prints the getters' results as one JSON line.
#include <cstdio>
#include <map>
#include <vector>

#include "sba/mono_sba_solver.h"

using namespace cuvslam;

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
  sba::MonoSBASolver solver(F, T, mi[0], mr[0], mr[1]);
  for (int t = 0; t < T; ++t) solver.addTrack(TrackId(t + 1), Vector3T(pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]));
  for (int k = 0; k < F; ++k) {
    std::vector<camera::Observation> obs;
    std::map<TrackId, bool> filtered;
    for (int o = 0; o < O; ++o)
      if (of[o] == k) {
        obs.emplace_back(CameraId(0), TrackId(ot[o] + 1), Vector2T(xy[2 * o], xy[2 * o + 1]), Matrix2T::Identity());
        filtered[TrackId(ot[o] + 1)] = false;
      }
    Isometry3T a;
    for (int i = 0; i < 16; ++i) a.data()[i] = poses[16 * k + i];
    solver.addFrame(FrameId(100 + k), con[k] != 0, a, filtered, obs);
  }
  const int it = solver.solve();
  const CameraMap cams = solver.getCameras();
  const Tracks3DMap tr = solver.getTracks3d();
  std::printf("{\"iterations\": %d, \"failed\": %s, \"residual\": %.9g, \"residual_initial\": %.9g, \"lambda\": %.9g, "
              "\"poses\": [", it, solver.getFailed() ? "true" : "false", solver.getResidual(),
              solver.getResidualInitial(), solver.getLambda());
  bool first = true;
  for (const auto& c : cams)
    for (int i = 0; i < 16; ++i) {
      std::printf("%s%.9g", first ? "" : ", ", c.second.data()[i]);
      first = false;
    }
  std::printf("], \"points\": [");
  first = true;
  for (const auto& p : tr)
    for (int d = 0; d < 3; ++d) {
      std::printf("%s%.9g", first ? "" : ", ", p.second(d));
      first = false;
    }
  const Matrix6T C = solver.getCovariance();
  std::printf("], \"cov\": [");
  for (int i = 0; i < 36; ++i) std::printf("%s%.9g", i ? ", " : "", C(i / 6, i % 6));
  std::printf("]}\n");
  return 0;
}
