// Host runtime of the batched mono SBA solve: the C ABI of include/cuvslam_sba_mono.h. Every window is packed on
// the host (sba_mono_pack.h), all windows' buffers share one device block in the handle's auxiliary workspace, one
// launch solves them (a workgroup per window, sba_mono.hip), and the results are read back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cuvslam_sba_mono.h"
#include "sba_internal.h"
#include "sba_mono_body.h"
#include "sba_mono_pack.h"

namespace sba_mono {
hipError_t launch_mono_solve(const MonoJob* jobs, int n, int lds_doubles, hipStream_t s);
}

using namespace sba_mono;

namespace {

constexpr int64_t kMonoLds = 18432;  // doubles of workgroup memory (144 KiB): 3 n^2 for n <= 78 (13 free frames)
constexpr int kMaxMonoFree = 64;  // free frames per window: reduced system of order <= 384

struct Slab {  // host staging of everything a batch uploads, with the device offset of each piece
  std::vector<char> bytes;
  size_t add(const void* src, size_t n) {
    const size_t o = bytes.size();
    bytes.resize(o + ((n + 255) & ~size_t(255)));
    if (n && src) std::memcpy(bytes.data() + o, src, n);
    return o;
  }
};

// Items of the sliced sums of a window: up to a workgroup's threads, and every frame or reduced-system block when
// there are more of those (then one slice each).
size_t part_items(const MonoPack& k) {
  return std::max<size_t>({size_t(kMonoThreads), size_t(k.nF), size_t(k.nFree) * size_t(k.nFree)});
}

}  // namespace

extern "C" {

int sba_mono_solve_batch(sba_handle* h, const sba_mono_window* windows, int32_t n, sba_mono_result* results) {
  using sba_internal::set_error;
  if (!h) return SBA_E_INVALID;
  if (n < 0 || (n > 0 && (!windows || !results))) return set_error(h, SBA_E_INVALID, "null windows or results");
  if (n == 0) return SBA_OK;
  const hipStream_t st = sba_internal::stream_of(h);
  if (hipError_t e = hipSetDevice(sba_internal::device_of(h)); e != hipSuccess)
    return set_error(h, SBA_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));

  // inputs into one slab; per-window work areas follow in the same device block
  std::vector<MonoPack> packs(n);
  Slab in;
  struct Offs {
    size_t con, fidx, pose, fro, fob, pts, tro, tob, ofr, otr, oxy, blk, bpr;
  };
  std::vector<Offs> offs(n);
  for (int i = 0; i < n; ++i) {
    const sba_mono_window& v = windows[i];
    if (v.num_frames < 0 || v.num_tracks < 0 || v.num_obs < 0 ||
        (v.num_frames && (!v.constrained || !v.frame_poses)) || (v.num_tracks && !v.points) ||
        (v.num_obs && (!v.obs_frame || !v.obs_track || !v.obs_xy)))
      return set_error(h, SBA_E_INVALID, "window " + std::to_string(i) + ": bad sizes or null buffers");
    if (!pack_window(v, packs[i]))
      return set_error(h, SBA_E_INVALID, "window " + std::to_string(i) + ": observation indexes out of range");
    if (packs[i].nFree > kMaxMonoFree)
      return set_error(h, SBA_E_CAPACITY, "window " + std::to_string(i) + " has " + std::to_string(packs[i].nFree) +
                                              " free frames; at most " + std::to_string(kMaxMonoFree));
    const MonoPack& k = packs[i];
    Offs& o = offs[i];
    o.con = in.add(k.constrained.data(), 4 * k.constrained.size());
    o.fidx = in.add(k.free_idx.data(), 4 * k.free_idx.size());
    o.pose = in.add(v.frame_poses, 64 * size_t(k.nF));
    o.fro = in.add(k.frame_obs.data(), 4 * k.frame_obs.size());
    o.fob = in.add(k.fobs.data(), 4 * k.fobs.size());
    o.pts = in.add(v.points, 12 * size_t(k.nT));
    o.tro = in.add(k.track_obs.data(), 4 * k.track_obs.size());
    o.tob = in.add(k.tobs.data(), 4 * k.tobs.size());
    o.ofr = in.add(k.oframe.data(), 4 * k.oframe.size());
    o.otr = in.add(k.otrack.data(), 4 * k.otrack.size());
    o.oxy = in.add(v.obs_xy, 8 * size_t(k.nO));
    o.blk = in.add(k.blk.data(), 4 * k.blk.size());
    o.bpr = in.add(k.bpair.data(), 4 * k.bpair.size());
  }
  // work areas (doubles unless noted), then every window's outputs in one region (read back with one copy):
  // scalars (8 doubles), covariance (36 doubles), poses (16 floats per frame), points (3 floats per track)
  std::vector<size_t> work(n), outp(n);
  size_t total = in.bytes.size();
  auto round = [](size_t b) { return (b + 255) & ~size_t(255); };
  constexpr size_t kResHead = 64 + 288;
  std::vector<size_t> wbytes(n);
  for (int i = 0; i < n; ++i) {
    const MonoPack& k = packs[i];
    const size_t F = k.nF, T = k.nT, O = k.nO, m = 6 * size_t(k.nFree);
    const size_t doubles = F * (6 * 5 + 36 * 2 + 12 + 9) + T * (3 * 4 + 9 * 2 + 4) + O * (2 + 1 + 12 + 6 + 18 + 18) +
                           4 * (m * m + 1) + 6 * (m + 1) + 2 * (m + 2) + 36 + part_items(k) * kPart + 64;
    wbytes[i] = round(8 * doubles + 4 * (T + 1) + 40 * 256);  // every array is rounded up to 256 bytes
    work[i] = total;
    total += wbytes[i];
  }
  const size_t res_base = total;
  for (int i = 0; i < n; ++i) {
    outp[i] = total - res_base;
    total += round(kResHead + 64 * size_t(packs[i].nF) + 12 * size_t(packs[i].nT));
  }
  const size_t res_bytes = total - res_base;
  const size_t jobs_off = total;
  total += round(sizeof(MonoJob) * size_t(n));
  char* base = nullptr;
  if (int rc = sba_internal::aux_workspace(h, total, &base); rc != SBA_OK) return rc;

  std::vector<MonoJob> jobs(n);
  for (int i = 0; i < n; ++i) {
    const MonoPack& k = packs[i];
    const Offs& o = offs[i];
    MonoWin& w = jobs[i].w;
    w = MonoWin{};
    w.nF = k.nF;
    w.nT = k.nT;
    w.nO = k.nO;
    w.nFree = k.nFree;
    w.max_iterations = windows[i].max_iterations;
    w.min_residual = windows[i].min_residual;
    w.min_residual_speed = windows[i].min_residual_speed;
    w.constrained = reinterpret_cast<const int32_t*>(base + o.con);
    w.free_idx = reinterpret_cast<const int32_t*>(base + o.fidx);
    w.pose_in = reinterpret_cast<const float*>(base + o.pose);
    w.frame_obs = reinterpret_cast<const int32_t*>(base + o.fro);
    w.fobs = reinterpret_cast<const int32_t*>(base + o.fob);
    w.pts_in = reinterpret_cast<const float*>(base + o.pts);
    w.track_obs = reinterpret_cast<const int32_t*>(base + o.tro);
    w.tobs = reinterpret_cast<const int32_t*>(base + o.tob);
    w.oframe = reinterpret_cast<const int32_t*>(base + o.ofr);
    w.otrack = reinterpret_cast<const int32_t*>(base + o.otr);
    w.oxy = reinterpret_cast<const float*>(base + o.oxy);
    w.blk = reinterpret_cast<const int32_t*>(base + o.blk);
    w.bpair = reinterpret_cast<const int2*>(base + o.bpr);
    const size_t F = k.nF, T = k.nT, O = k.nO, m = 6 * size_t(k.nFree);
    char* p = base + work[i];
    auto take = [&](size_t count) {
      double* r = reinterpret_cast<double*>(p);
      p += round(8 * std::max<size_t>(count, 1));
      return r;
    };
    w.fa = take(6 * F); w.faD = take(6 * F); w.fda = take(6 * F); w.ferrA = take(6 * F); w.feps = take(6 * F);
    w.fUo = take(36 * F); w.fUa = take(36 * F); w.fM = take(12 * F); w.fRn = take(9 * F);
    w.tb = take(3 * T); w.tbD = take(3 * T); w.tdb = take(3 * T); w.terrB = take(3 * T); w.tV = take(9 * T);
    w.tinvVa = take(9 * T); w.tinfo = take(4 * T);
    w.oerr = take(2 * O); w.ow = take(O); w.oA = take(12 * O); w.oB = take(6 * O); w.oW = take(18 * O);
    w.oY = take(18 * O);
    w.S = take(m * m + 1); w.Sb = take(m + 1); w.Q = take(m * m + 1); w.E = take(m * m + 1); w.E2 = take(m * m + 1); w.ev = take(m + 1);
    w.x = take(m + 1); w.xt = take(m + 1); w.bd = take(m + 1); w.dv = take(m + 1); w.rot = take(2 * (m + 2));
    w.prep = take(36);
    w.part = take(part_items(k) * kPart);
    w.tskip = reinterpret_cast<int32_t*>(p);
    p += round(4 * (T + 1));
    if (p > base + work[i] + wbytes[i]) return set_error(h, SBA_E_INVALID, "mono work area overflow");
    char* r = base + res_base + outp[i];
    w.out_scalars = reinterpret_cast<double*>(r);
    w.cov = reinterpret_cast<double*>(r + 64);
    jobs[i].poses_out = reinterpret_cast<float*>(r + kResHead);
    jobs[i].pts_out = reinterpret_cast<float*>(r + kResHead + 64 * F);
  }
  hipError_t e = hipMemcpyAsync(base, in.bytes.data(), in.bytes.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(base + jobs_off, jobs.data(), sizeof(MonoJob) * size_t(n), hipMemcpyHostToDevice, st);
  // workgroup memory for the eigen-decomposition of the largest reduced system, up to kMonoLds
  int64_t lds = 0;
  for (int i = 0; i < n; ++i) {
    const int64_t m = 6 * packs[i].nFree;
    lds = std::max<int64_t>(lds, 3 * m * m + 2 * (m + 2));  // E, its second buffer, Q, a step's rotations
  }
  if (lds > kMonoLds) lds = 0;  // the largest does not fit: every window keeps its matrices in its work area
  if (e == hipSuccess)
    e = launch_mono_solve(reinterpret_cast<const MonoJob*>(base + jobs_off), n, static_cast<int>(lds), st);
  if (e != hipSuccess) return set_error(h, SBA_E_DEVICE, std::string("mono solve: ") + hipGetErrorString(e));
  std::vector<char> host_res(res_bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(host_res.data(), base + res_base, res_bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_error(h, SBA_E_DEVICE, std::string("mono results: ") + hipGetErrorString(e));
  double sc[8], cov[36];
  for (int i = 0; i < n; ++i) {
    const MonoPack& k = packs[i];
    sba_mono_result& r = results[i];
    const char* src = host_res.data() + outp[i];
    std::memcpy(sc, src, 64);
    std::memcpy(cov, src + 64, 288);
    if (r.frame_poses && k.nF) std::memcpy(r.frame_poses, src + kResHead, 64 * size_t(k.nF));
    if (r.points && k.nT) std::memcpy(r.points, src + kResHead + 64 * size_t(k.nF), 12 * size_t(k.nT));
    for (int q = 0; q < 36; ++q) r.covariance[q] = float(cov[q]);
    r.residual = sc[0];
    r.residual_initial = sc[1];
    r.lambda = sc[2];
    r.iterations = int32_t(sc[3]);
    r.failed = int32_t(sc[4]);
    r.feasible = int32_t(sc[5]);
    r.status = r.feasible ? SBA_OK : SBA_E_INFEASIBLE;
  }
  return SBA_OK;
}

}  // extern "C"
