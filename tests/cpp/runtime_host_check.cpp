// CPU check of the host runtime (cuvslam_amd/csrc/sba_runtime.cpp) with no GPU: the runtime is linked against
// stand-ins of the HIP calls it makes, whose "device" memory is host memory, and of the kernel launches, which run
// nothing but check the batch they are handed: every stream a kernel reads or writes is bound, lies inside a live
// allocation with room for what the kernels index, and does not overlap another stream. The entry points are then
// driven the way the tests drive them (one window, the handle's workspace growing and shrinking, the fixed scratch,
// uneven batches, multi-camera windows whose pairs hold several observations, a mapped window).
// Prints one line per case; exits non-zero on a failed check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../cuvslam_amd/csrc/sba_device.h"
#include "../../include/cuvslam_sba.h"

using namespace sba_dev;

static int g_fail = 0;
static std::string g_case;
#define CHECK(cond, what)                                                                   \
  do {                                                                                      \
    if (!(cond)) {                                                                          \
      std::printf("  FAILED [%s] %s (%s:%d)\n", g_case.c_str(), what, __FILE__, __LINE__); \
      ++g_fail;                                                                             \
    }                                                                                       \
  } while (0)

// ---- "device" memory: tracked host allocations ----------------------------------------------------------------
static std::map<const char*, size_t> g_allocs;  // base -> bytes
static int g_mallocs = 0, g_frees = 0;
static bool g_forbid_alloc = false;

static bool inside(const void* p, size_t bytes) {
  const char* c = static_cast<const char*>(p);
  auto it = g_allocs.upper_bound(c);
  if (it == g_allocs.begin()) return false;
  --it;
  return c >= it->first && c + bytes <= it->first + it->second;
}

extern "C" {
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* s) { *s = reinterpret_cast<hipStream_t>(0x1); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t bytes) {
  CHECK(!g_forbid_alloc, "hipMalloc called where no allocation is allowed");
  *p = std::calloc(1, bytes ? bytes : 1);
  g_allocs[static_cast<char*>(*p)] = bytes;
  ++g_mallocs;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  CHECK(!g_forbid_alloc, "hipFree called where no allocation is allowed");
  if (!p) return hipSuccess;
  CHECK(g_allocs.count(static_cast<char*>(p)) == 1, "hipFree of an unknown pointer");
  g_allocs.erase(static_cast<char*>(p));
  std::free(p);
  ++g_frees;
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
  if (k == hipMemcpyHostToDevice) CHECK(inside(d, n), "H2D copy outside an allocation");
  if (k == hipMemcpyDeviceToHost) CHECK(inside(s, n), "D2H copy outside an allocation");
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemset(void* p, int v, size_t n) {
  CHECK(inside(p, n), "memset outside an allocation");
  std::memset(p, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { return hipMemset(p, v, n); }
hipError_t hipEventCreate(hipEvent_t* e) { *e = reinterpret_cast<hipEvent_t>(0x2); return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 1.f; return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = reinterpret_cast<hipGraph_t>(0x3); return hipSuccess; }
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  *e = reinterpret_cast<hipGraphExec_t>(0x4);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* prop, int) {
  std::memset(prop, 0, sizeof(*prop));
  std::strcpy(prop->gcnArchName, "gfx950:sramecc+:xnack-");
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "stub"; }
hipError_t hipGetLastError() { return hipSuccess; }
}

// ---- the checks the kernel stand-ins make ---------------------------------------------------------------------
struct Range {
  const char* name;
  const void* p;
  size_t bytes;
};

static void check_ranges(std::vector<Range> rs) {
  for (const auto& r : rs) {
    CHECK(r.p != nullptr, (std::string("stream not bound: ") + r.name).c_str());
    if (r.p) CHECK(inside(r.p, r.bytes), (std::string("stream outside its allocation: ") + r.name).c_str());
  }
  // distinct streams must not overlap (the pair streams may alias the observation streams)
  std::sort(rs.begin(), rs.end(), [](const Range& a, const Range& b) { return a.p < b.p; });
  for (size_t i = 1; i < rs.size(); ++i) {
    if (rs[i].p == rs[i - 1].p) continue;
    const char* end = static_cast<const char*>(rs[i - 1].p) + rs[i - 1].bytes;
    CHECK(end <= static_cast<const char*>(rs[i].p),
          (std::string("streams overlap: ") + rs[i - 1].name + " and " + rs[i].name).c_str());
  }
}

static void check_batch(const Batch& b) {
  const int nw = b.n_windows;
  CHECK(nw > 0, "no windows");
  if (!b.win || !inside(b.win, sizeof(WindowDesc) * nw)) {
    CHECK(false, "window table not bound");
    return;
  }
  int64_t nobs = 0, npairs = 0, part = 0, sred = 0, dc = 0, a = 0, big = 0, ue = 0;
  for (int w = 0; w < nw; ++w) {
    const WindowDesc& W = b.win[w];
    nobs = std::max<int64_t>(nobs, W.obs_off + W.no);
    sred = std::max<int64_t>(sred, W.sred_off + W.rec);
    dc = std::max<int64_t>(dc, W.dc_off + 6 * W.ncf);
    if (W.mapped) {
      const int64_t n = 6 * W.ncf;
      a = std::max<int64_t>(a, W.a_off + n * n + 2 * n + ((n + 63) / 64) * 64 * 64);
      big = std::max<int64_t>(big, W.big_off + n * n);
      ue = std::max<int64_t>(ue, W.ue_off + 33 * W.ncf);
    }
    CHECK(W.nbchunks >= 0 && W.bchunk_begin >= 0 && W.bchunk_begin + W.nbchunks <= b.n_bchunks, "window's cost slots");
  }
  int64_t ntiles = 0, nbtiles = 0;
  for (int c = 0; c < b.n_chunks; ++c) {
    const Chunk& ch = b.chunks[c];
    ntiles = std::max<int64_t>(ntiles, ch.tile_end);
    const int64_t rec = 256 * (ch.nt * (ch.nt + 1) / 2) + 33 * ch.nloc;
    part = std::max<int64_t>(part, ch.part_off + rec);
    CHECK(ch.window >= 0 && ch.window < nw, "chunk window");
  }
  for (int c = 0; c < b.n_bchunks; ++c) nbtiles = std::max<int64_t>(nbtiles, b.bchunks[c].tile_end);
  for (int64_t t = 0; t < ntiles; ++t) {
    const Tile& T = b.tiles[t];
    npairs = std::max<int64_t>(npairs, T.pair_end);
    CHECK(T.point_end <= b.n_points && T.pair_begin <= T.pair_end, "tile ranges");
  }
  for (int64_t t = 0; t < nbtiles; ++t) npairs = std::max<int64_t>(npairs, b.btiles[t].pair_end);
  for (int64_t q = 0; q < npairs && b.pair_rec; ++q) CHECK(b.pair_rec[q].y <= nobs && b.pair_rec[q].x < b.pair_rec[q].y, "pair record");
  std::vector<Range> rs = {
      {"win", b.win, sizeof(WindowDesc) * nw}, {"st", b.st, sizeof(WindowState) * nw},
      {"tiles", b.tiles, sizeof(Tile) * ntiles}, {"chunks", b.chunks, sizeof(Chunk) * b.n_chunks},
      {"btiles", b.btiles, sizeof(Tile) * nbtiles}, {"bchunks", b.bchunks, sizeof(Chunk) * b.n_bchunks},
      {"red_first", b.red_first, 4 * static_cast<size_t>(nw + 1)}, {"obs_xy", b.obs_xy, 8 * nobs},
      {"obs_info", b.obs_info, 16 * nobs}, {"pair_obs", b.pair_obs, 4 * static_cast<size_t>(npairs + 1)},
      {"pair_pose", b.pair_pose, 4 * npairs}, {"pair_rec", b.pair_rec, 16 * npairs},
      {"pair_xy", b.pair_xy, 8 * npairs}, {"pair_info", b.pair_info, 16 * npairs},
      {"point_pair", b.point_pair, 4 * static_cast<size_t>(b.n_points + 1)}, {"point_nfree", b.point_nfree, 4 * static_cast<size_t>(b.n_points)},
      {"pts0", b.pts0, 16 * static_cast<size_t>(b.n_points)}, {"pts1", b.pts1, 16 * static_cast<size_t>(b.n_points)},
      {"pts_init", b.pts_init, 16 * static_cast<size_t>(b.n_points)}, {"paux", b.paux, 48 * static_cast<size_t>(b.n_points)},
      {"poses0", b.poses0, 48 * static_cast<size_t>(b.n_poses)}, {"poses1", b.poses1, 48 * static_cast<size_t>(b.n_poses)},
      {"poses_init", b.poses_init, 48 * static_cast<size_t>(b.n_poses)}, {"cams", b.cams, 48 * static_cast<size_t>(b.n_cams)},
      {"part", b.part, 4 * part}, {"sred", b.sred, 4 * sred}, {"dc", b.dc, 4 * dc}, {"dc64", b.dc64, 8 * dc},
      {"out_pts", b.out_pts, 16 * static_cast<size_t>(b.n_points)}, {"out_poses", b.out_poses, 48 * static_cast<size_t>(b.n_poses)},
      {"trace", b.trace, sizeof(TraceEntry) * kTrace * nw}, {"bpart", b.bpart, sizeof(CostPart) * b.n_bchunks},
      {"A", b.A, 8 * a}, {"big", b.big, 4 * big}, {"big_ue", b.big_ue, 4 * ue}, {"chunk_map", b.chunk_map, 4}};
  // the pair streams alias the observation streams when every pair has one observation
  if (b.pair_xy == b.obs_xy) {
    CHECK(npairs == nobs, "pair streams alias observation streams of another length");
    rs.erase(std::remove_if(rs.begin(), rs.end(), [](const Range& r) { return !std::strcmp(r.name, "pair_xy") || !std::strcmp(r.name, "pair_info"); }), rs.end());
  }
  check_ranges(rs);
}

// The packed streams of the last batch whose solve was enqueued (copied at its first launch).
static std::vector<std::vector<char>> g_packed;
static void snapshot(const Batch& b) {
  auto grab = [](const void* p, size_t bytes) {
    const char* c = static_cast<const char*>(p);
    return std::vector<char>(c, c + bytes);
  };
  int64_t nobs = 0, npairs = 0, ntiles = 0, nbtiles = 0;
  for (int w = 0; w < b.n_windows; ++w) nobs = std::max<int64_t>(nobs, b.win[w].obs_off + b.win[w].no);
  for (int c = 0; c < b.n_chunks; ++c) ntiles = std::max<int64_t>(ntiles, b.chunks[c].tile_end);
  for (int c = 0; c < b.n_bchunks; ++c) nbtiles = std::max<int64_t>(nbtiles, b.bchunks[c].tile_end);
  for (int64_t t = 0; t < ntiles; ++t) npairs = std::max<int64_t>(npairs, b.tiles[t].pair_end);
  g_packed = {grab(b.win, sizeof(WindowDesc) * b.n_windows), grab(b.tiles, sizeof(Tile) * ntiles),
              grab(b.chunks, sizeof(Chunk) * b.n_chunks), grab(b.btiles, sizeof(Tile) * nbtiles),
              grab(b.bchunks, sizeof(Chunk) * b.n_bchunks), grab(b.obs_xy, 8 * nobs), grab(b.obs_info, 16 * nobs),
              grab(b.pair_obs, 4 * (npairs + 1)), grab(b.pair_pose, 4 * npairs), grab(b.pair_rec, 16 * npairs),
              grab(b.pair_xy, 8 * npairs), grab(b.pair_info, 16 * npairs),
              grab(b.point_pair, 4 * (b.n_points + 1)), grab(b.point_nfree, 4 * b.n_points),
              grab(b.pts_init, 16 * b.n_points), grab(b.poses_init, 48 * b.n_poses)};
}

namespace sba_dev {
static int g_launches = 0;
hipError_t launch_reset(const Batch& b, hipStream_t) {
  ++g_launches;
  check_batch(b);
  snapshot(b);
  return hipSuccess;
}
hipError_t launch_schur(const Batch& b, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
hipError_t launch_reduce(const Batch& b, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
hipError_t launch_factor(const Batch& b, double*, hipStream_t, double*) {
  ++g_launches;
  CHECK(b.win && b.st && b.sred && b.dc && b.dc64 && b.poses0 && b.poses1, "factor streams bound");
  return hipSuccess;
}
hipError_t launch_backsub(const Batch& b, bool, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
hipError_t launch_decide(const Batch& b, bool, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
hipError_t launch_gather(const Batch& b, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
hipError_t launch_large_zero(const Batch& b, int, hipStream_t) { check_batch(b); return hipSuccess; }
hipError_t launch_large_scatter(const Batch& b, int, int, hipStream_t) { check_batch(b); return hipSuccess; }
hipError_t launch_large_factor(const Batch& b, int, int, hipStream_t) { check_batch(b); return hipSuccess; }
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t) {
  CHECK(inside(p, bytes), "fill outside an allocation");
  CHECK(reinterpret_cast<uintptr_t>(p) % 16 == 0, "fill start not 16-byte aligned");
  std::memset(p, 0, bytes);
  return hipSuccess;
}
hipError_t launch_adj_obs(const Batch&, int, const double*, const double*, const double*, double*, double*, const double*,
                          double*, double*, const int32_t*, int32_t*, hipStream_t) { return hipSuccess; }
hipError_t launch_adj_state(const Batch&, double*, double*, const double*, int, hipStream_t) { return hipSuccess; }
hipError_t launch_adj_chol(const double*, int, const double*, double*, const int32_t*, hipStream_t) { return hipSuccess; }
hipError_t launch_adj_vec(const Batch&, double*, double*, double*, double*, double*, double*, double*, double*, double*,
                          double*, int32_t*, double, int, int, int, hipStream_t) { return hipSuccess; }
hipError_t launch_winsorize_select(const Batch&, int, int, int, double*, int32_t*, int32_t*, uint8_t*, int32_t*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_winsorize_apply(const Batch&, int, int, int, const uint8_t*, hipStream_t) { return hipSuccess; }
hipError_t launch_copy_calibrate(float4*, const float4*, int64_t, hipStream_t) { return hipSuccess; }
hipError_t launch_mfma_layout(const float*, const float*, float*, hipStream_t) { return hipSuccess; }
hipError_t set_kernel_attributes() { return hipSuccess; }
static void* g_scratch = nullptr;
constexpr size_t kStubScratch = size_t(64) << 20;
int scratch_setup(int, void** ptr, size_t* bytes) {
  if (!g_scratch) {
    g_scratch = std::calloc(1, kStubScratch);
    g_allocs[static_cast<char*>(g_scratch)] = kStubScratch;
  }
  *ptr = g_scratch;
  *bytes = kStubScratch;
  return 0;
}
}  // namespace sba_dev

// ---- driving the entry points ---------------------------------------------------------------------------------
struct Scene {
  sba_scene* s = nullptr;
  sba_problem_view v{};
  std::vector<float> pts, poses;
  ~Scene() { sba_scene_free(s); }
};

static void make_scene(Scene& sc, int kf, int pts, uint64_t seed, int cams = 1, int iters = 3) {
  sba_scene_config c;
  sba_scene_default_config(&c);
  c.num_keyframes = kf;
  c.num_points = pts;
  c.seed = seed;
  c.num_fixed_keyframes = 2;
  c.sample_observations = 1;
  (void)cams;
  sba_scene_generate(&c, &sc.s);
  sba_scene_view(sc.s, &sc.v);
  sc.v.max_iterations = iters;
  sc.pts.resize(3 * sc.v.num_points);
  sc.poses.resize(16 * sc.v.num_poses);
}

static void make_loop(Scene& sc, int kf, int pts) {
  sba_loop_config c;
  sba_loop_default_config(&c);
  c.num_keyframes = kf;
  c.num_points = pts;
  c.loop_length_m = 0.2 * kf;
  sba_scene_generate_loop(&c, &sc.s);
  sba_scene_view(sc.s, &sc.v);
  sc.v.max_iterations = 2;
  sc.pts.resize(3 * sc.v.num_points);
  sc.poses.resize(16 * sc.v.num_poses);
}

static sba_result result_for(Scene& sc) {
  sba_result r{};
  r.points = sc.pts.data();
  r.rig_from_world = sc.poses.data();
  return r;
}

int main() {
  sba_options o;
  sba_default_options(&o);
  o.loss = SBA_LOSS_HUBER;
  sba_handle* h = nullptr;
  g_case = "create";
  CHECK(sba_create(&o, &h) == SBA_OK, "sba_create");

  g_case = "solve: workspace grows, shrinks, grows";
  std::printf("%s\n", g_case.c_str());
  for (int k : {0, 1, 2}) {
    Scene sc;
    make_scene(sc, k == 1 ? 5 : 12 + 8 * k, k == 1 ? 200 : 1500 + 1000 * k, 10 + k);
    sba_result r = result_for(sc);
    const int rc = sba_solve(h, &sc.v, &r);
    CHECK(rc == SBA_OK, (std::string("sba_solve: ") + sba_last_error(h)).c_str());
  }

  g_case = "solve_noalloc: no allocation, the capacity error";
  std::printf("%s\n", g_case.c_str());
  {
    const int mallocs = g_mallocs, frees = g_frees;
    g_forbid_alloc = true;
    Scene a, b;
    make_scene(a, 14, 3000, 21);
    make_scene(b, 6, 600, 22);
    sba_result ra = result_for(a), rb = result_for(b);
    CHECK(sba_solve_noalloc(h, &a.v, &ra) == SBA_OK, sba_last_error(h));
    CHECK(sba_solve_noalloc(h, &b.v, &rb) == SBA_OK, sba_last_error(h));
    Scene big;
    make_scene(big, 40, 40000, 23);
    sba_result rbig = result_for(big);
    CHECK(sba_solve_noalloc(h, &big.v, &rbig) == SBA_E_CAPACITY, "capacity error");
    g_forbid_alloc = false;
    CHECK(g_mallocs == mallocs && g_frees == frees, "no allocation or free in sba_solve_noalloc");
  }

  g_case = "batch: uneven windows, graph";
  std::printf("%s\n", g_case.c_str());
  {
    std::vector<Scene> scs(6);
    std::vector<sba_problem_view> vs;
    for (int k = 0; k < 6; ++k) {
      make_scene(scs[k], 5 + 7 * k, 300 + 400 * k, 30 + k);
      vs.push_back(scs[k].v);
    }
    sba_batch* b = nullptr;
    CHECK(sba_batch_create(h, vs.data(), 6, &b) == SBA_OK, sba_last_error(h));
    CHECK(sba_batch_run(h, b) == SBA_OK, "run");
    std::vector<sba_result> rs(6);
    for (int k = 0; k < 6; ++k) rs[k] = result_for(scs[k]);
    CHECK(sba_batch_read(h, b, rs.data()) == SBA_OK, "read");
    float ms[5];
    int32_t n[5];
    CHECK(sba_batch_kernel_times(h, b, ms, n) == SBA_OK, "kernel times");
    double rho[64], cost[64];
    float lam[64];
    int32_t flags[64], nt = 0;
    CHECK(sba_batch_trace(h, b, 3, 64, rho, cost, lam, flags, &nt) == SBA_OK, "trace");
    sba_batch_destroy(h, b);
  }

  g_case = "batch: multi-camera map window (pairs with several observations, mapped factor)";
  std::printf("%s\n", g_case.c_str());
  {
    Scene loop;
    make_loop(loop, 48, 3000);
    Scene small;
    make_scene(small, 8, 500, 41);
    sba_problem_view vs[2] = {small.v, loop.v};
    sba_batch* b = nullptr;
    CHECK(sba_batch_create(h, vs, 2, &b) == SBA_OK, sba_last_error(h));
    CHECK(sba_batch_run(h, b) == SBA_OK, "run");
    sba_batch_destroy(h, b);
    sba_result r = result_for(loop);
    CHECK(sba_solve(h, &loop.v, &r) == SBA_OK, sba_last_error(h));
  }

  g_case = "packing on 1 and on 8 host threads gives the same streams";
  std::printf("%s\n", g_case.c_str());
  {
    Scene big, rig;
    make_scene(big, 30, 5000, 51);
    make_loop(rig, 48, 30000);  // pairs with several observations
    for (Scene* sc : {&big, &rig}) {
      CHECK(sc->v.num_observations >= (1 << 16), "a window large enough for the threaded packing");
      std::vector<std::vector<char>> packed[2];
      for (int threads : {1, 8}) {
        sba_options o1 = o;
        o1.pack_threads = threads;
        sba_handle* h1 = nullptr;
        CHECK(sba_create(&o1, &h1) == SBA_OK, "sba_create");
        sba_batch* b = nullptr;
        CHECK(sba_batch_create(h1, &sc->v, 1, &b) == SBA_OK, sba_last_error(h1));
        packed[threads == 8] = g_packed;
        sba_batch_destroy(h1, b);
        sba_destroy(h1);
      }
      CHECK(packed[0].size() == packed[1].size() && !packed[0].empty(), "stream count");
      for (size_t k = 0; k < packed[0].size() && k < packed[1].size(); ++k)
        CHECK(packed[0][k] == packed[1][k], (std::string("stream ") + std::to_string(k) + " differs").c_str());
    }
  }

  g_case = "debug step";
  std::printf("%s\n", g_case.c_str());
  {
    Scene sc;
    make_scene(sc, 6, 300, 3);
    sba_batch* b = nullptr;
    CHECK(sba_batch_create(h, &sc.v, 1, &b) == SBA_OK, sba_last_error(h));
    double sc8[8];
    sba_batch_debug_step(h, b, 1e-3, nullptr, nullptr, nullptr, nullptr, sc8);
    sba_batch_destroy(h, b);
  }
  sba_destroy(h);
  g_case = "leaks";
  // the fake scratch stays (it lives as long as the process); everything else is freed
  CHECK(g_allocs.size() == 1, "every allocation freed");
  std::printf("%s (%d failed checks, %d launches)\n", g_fail ? "FAILED" : "OK", g_fail, sba_dev::g_launches);
  return g_fail ? 1 : 0;
}
