// Host packing of one mono SBA window (shared by the runtime and the CPU check): the observation orders and the
// pair lists of the reduced system that the solver body (sba_mono_body.h) reads. The orders follow how
// MonoSBASolver stores a window (mono_sba_solver.cpp:47-120): frames are added in order, so a track's observations
// are in frame order, and a frame's contributions accumulate in track order.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/cuvslam_sba_mono.h"

namespace sba_mono {

struct MonoPack {
  int32_t nF = 0, nT = 0, nO = 0, nFree = 0;
  std::vector<int32_t> constrained, free_idx, frame_obs, fobs, track_obs, tobs, oframe, otrack, blk;
  std::vector<int32_t> bpair;  // pairs (observation in j, observation in k), two ints each
};

// Returns false when an index is out of range.
inline bool pack_window(const sba_mono_window& v, MonoPack& k) {
  k.nF = v.num_frames;
  k.nT = v.num_tracks;
  k.nO = v.num_obs;
  k.constrained.assign(v.constrained, v.constrained + k.nF);
  k.free_idx.assign(k.nF, -1);
  k.nFree = 0;
  for (int f = 0; f < k.nF; ++f)
    if (!k.constrained[f]) k.free_idx[f] = k.nFree++;
  k.oframe.assign(v.obs_frame, v.obs_frame + k.nO);
  k.otrack.assign(v.obs_track, v.obs_track + k.nO);
  for (int o = 0; o < k.nO; ++o)
    if (k.oframe[o] < 0 || k.oframe[o] >= k.nF || k.otrack[o] < 0 || k.otrack[o] >= k.nT) return false;
  std::vector<int32_t> idx(k.nO);
  std::iota(idx.begin(), idx.end(), 0);
  k.tobs = idx;
  std::stable_sort(k.tobs.begin(), k.tobs.end(), [&](int a, int b) {
    return k.otrack[a] != k.otrack[b] ? k.otrack[a] < k.otrack[b] : k.oframe[a] < k.oframe[b];
  });
  k.fobs = idx;
  std::stable_sort(k.fobs.begin(), k.fobs.end(), [&](int a, int b) {
    return k.oframe[a] != k.oframe[b] ? k.oframe[a] < k.oframe[b] : k.otrack[a] < k.otrack[b];
  });
  k.track_obs.assign(k.nT + 1, 0);
  for (int o = 0; o < k.nO; ++o) ++k.track_obs[k.otrack[o] + 1];
  for (int t = 0; t < k.nT; ++t) k.track_obs[t + 1] += k.track_obs[t];
  k.frame_obs.assign(k.nF + 1, 0);
  for (int o = 0; o < k.nO; ++o) ++k.frame_obs[k.oframe[o] + 1];
  for (int f = 0; f < k.nF; ++f) k.frame_obs[f + 1] += k.frame_obs[f];
  // ordered pairs (j, k) of every track's free frames, per block in track order
  const int nb = k.nFree * k.nFree;
  k.blk.assign(nb + 1, 0);
  for (int t = 0; t < k.nT; ++t)
    for (int a = k.track_obs[t]; a < k.track_obs[t + 1]; ++a) {
      const int j = k.free_idx[k.oframe[k.tobs[a]]];
      if (j < 0) continue;
      for (int b = k.track_obs[t]; b < k.track_obs[t + 1]; ++b) {
        const int q = k.free_idx[k.oframe[k.tobs[b]]];
        if (q >= 0) ++k.blk[j * k.nFree + q + 1];
      }
    }
  for (int b = 0; b < nb; ++b) k.blk[b + 1] += k.blk[b];
  k.bpair.assign(2 * size_t(std::max(k.blk[nb], 1)), 0);
  std::vector<int32_t> fill(k.blk.begin(), k.blk.end() - 1);
  for (int t = 0; t < k.nT; ++t)
    for (int a = k.track_obs[t]; a < k.track_obs[t + 1]; ++a) {
      const int oj = k.tobs[a], j = k.free_idx[k.oframe[oj]];
      if (j < 0) continue;
      for (int b = k.track_obs[t]; b < k.track_obs[t + 1]; ++b) {
        const int ok = k.tobs[b], q = k.free_idx[k.oframe[ok]];
        if (q < 0) continue;
        const int slot = fill[j * k.nFree + q]++;
        k.bpair[2 * size_t(slot)] = oj;
        k.bpair[2 * size_t(slot) + 1] = ok;
      }
    }
  return true;
}

}  // namespace sba_mono
