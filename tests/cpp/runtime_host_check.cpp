// CPU check of the host runtime (cuvslam_amd/csrc/sba_runtime.cpp) with no GPU: the runtime is linked against
// stand-ins of the HIP calls it makes, whose "device" memory is host memory, and of the kernel launches, which run
// nothing but check the batch they are handed: every stream a kernel reads or writes is bound, lies inside a live
// allocation with room for what the kernels index, and does not overlap another stream. The entry points are then
// driven the way the tests drive them (one window, the handle's workspace growing and shrinking, the fixed scratch,
// uneven batches, multi-camera windows whose pairs hold several observations, a mapped window).
// Prints one line per case; exits non-zero on a failed check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../cuvslam_amd/csrc/sba_device.h"
#include "../../cuvslam_amd/csrc/sba_imu_device.h"
#include "../../include/cuvslam_sba.h"
#include "../../include/cuvslam_sba_imu.h"

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
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned int) {
  *p = std::malloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  std::free(p);
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

static int g_wide_points = 0;  // wide points seen in checked batches

static void check_batch(const Batch& b) {
  const int nw = b.n_windows;
  CHECK(nw > 0, "no windows");
  if (!b.win || !inside(b.win, sizeof(WindowDesc) * nw)) {
    CHECK(false, "window table not bound");
    return;
  }
  int64_t nobs = 0, npairs = 0, part = 0, sred = 0, dc = 0, a = 0, gf = 1, gc = 1, cmap = 1;
  for (int w = 0; w < nw; ++w) {
    const WindowDesc& W = b.win[w];
    nobs = std::max<int64_t>(nobs, W.obs_off + W.no);
    sred = std::max<int64_t>(sred, W.sred_off + W.rec);
    dc = std::max<int64_t>(dc, W.dc_off + 6 * W.ncf);
    if (W.mapped) {
      const int64_t n = 6 * W.ncf;
      a = std::max<int64_t>(a, W.a_off + n * n + 2 * n + ((n + 63) / 64) * 64 * 64);
      gf = std::max<int64_t>(gf, W.gfirst_off + static_cast<int64_t>(W.ncf) * W.ncf + 1);
    }
    CHECK(W.nbchunks >= 0 && W.bchunk_begin >= 0 && W.bchunk_begin + W.nbchunks <= b.n_bchunks, "window's cost slots");
  }
  int64_t ntiles = 0, nbtiles = 0;
  for (int c = 0; c < b.n_chunks; ++c) {
    const Chunk& ch = b.chunks[c];
    ntiles = std::max<int64_t>(ntiles, ch.tile_end);
    const int64_t rec = b.win[ch.window].mapped ? 36 * (ch.nloc * (ch.nloc + 1) / 2) + 33 * ch.nloc
                                                : 256 * (ch.nt * (ch.nt + 1) / 2) + 33 * ch.nloc;
    part = std::max<int64_t>(part, ch.part_off + rec);
    CHECK(ch.window >= 0 && ch.window < nw, "chunk window");
  }
  for (int c = 0; c < b.n_chunks; ++c)
    if (b.win[b.chunks[c].window].mapped) cmap = std::max<int64_t>(cmap, b.chunks[c].map_off + b.chunks[c].nloc);
  // the gather plan of each mapped window: block (a, b) lists its contributions in increasing chunk order, each
  // naming a chunk of the window that stages a and b at the given local indices; every ordinary chunk that stages
  // both contributes, and the tiles of a wide point (a point seen from more than kMaxFreePoses free key frames)
  // contribute every block between two of its free key frames exactly once
  for (int w = 0; w < nw; ++w) {
    const WindowDesc& W = b.win[w];
    if (!W.mapped) continue;
    const int ncf = W.ncf;
    const int32_t* f = b.gfirst + W.gfirst_off;
    gc = std::max<int64_t>(gc, f[static_cast<int64_t>(ncf) * ncf]);
    std::vector<int> ordinary(static_cast<size_t>(ncf) * ncf, 0);
    std::map<int, std::set<int>> wide_kfs;  // wide point -> the free key frames its tiles stage
    for (int c = W.chunk_begin; c < W.chunk_begin + W.nchunks; ++c) {
      const Chunk& ch = b.chunks[c];
      CHECK(ch.window == w && ch.nloc <= kMaxFreePoses, "mapped chunk");
      const Tile& T = b.tiles[ch.tile_begin];
      if (T.virt) {
        CHECK(ch.tile_end == ch.tile_begin + 1 && T.point_end == T.point_begin + 1, "a wide point's chunk is one tile");
        for (int l = 0; l < ch.nloc; ++l) wide_kfs[T.point_begin].insert(b.chunk_map[ch.map_off + l]);
        continue;
      }
      for (int la = 0; la < ch.nloc; ++la)
        for (int lb = la; lb < ch.nloc; ++lb) {
          const int ka = b.chunk_map[ch.map_off + la], kb = b.chunk_map[ch.map_off + lb];
          CHECK(ka < kb || la == lb, "chunk map increasing");
          ++ordinary[ka * ncf + kb];
        }
    }
    std::map<std::pair<int, int>, int> wide_count;  // (wide point, block) -> contributions
    for (int ka = 0; ka < ncf; ++ka)
      for (int kb = 0; kb < ncf; ++kb) {
        const int blk = ka * ncf + kb;
        int n_ord = 0, last = -1;
        for (int k = f[blk]; k < f[blk + 1]; ++k) {
          const int c = b.gcon[k].x, la = b.gcon[k].y & 255, lb = (b.gcon[k].y >> 8) & 255;
          CHECK(c > last && c >= W.chunk_begin && c < W.chunk_begin + W.nchunks, "gather plan: chunk order");
          last = c;
          if (c < W.chunk_begin || c >= W.chunk_begin + W.nchunks) break;
          const Chunk& ch = b.chunks[c];
          CHECK(la <= lb && lb < ch.nloc && b.chunk_map[ch.map_off + la] == ka && b.chunk_map[ch.map_off + lb] == kb,
                "gather plan: local indices");
          CHECK(((b.gcon[k].y >> 16) & 255) == ch.nt && (b.gcon[k].y >> 24) == ch.nloc &&
                ((static_cast<int64_t>(b.gcon[k].w) << 32) | static_cast<uint32_t>(b.gcon[k].z)) == ch.part_off,
                "gather plan: the chunk's record"); 
          const Tile& T = b.tiles[ch.tile_begin];
          if (T.virt) ++wide_count[{T.point_begin, blk}];
          else ++n_ord;
        }
        CHECK(n_ord == ordinary[blk], "gather plan: the ordinary chunks that stage a block");
      }
    for (const auto& [pt, kfs] : wide_kfs) {
      CHECK(kfs.size() > static_cast<size_t>(kMaxFreePoses), "a wide point sees more than kMaxFreePoses key frames");
      for (int ka : kfs)
        for (int kb : kfs)
          if (ka <= kb) {
            auto it = wide_count.find({pt, ka * ncf + kb});
            CHECK(it != wide_count.end() && it->second == 1, "gather plan: a wide point's block summed once");
          }
    }
    size_t n_wide = 0;
    for (const auto& [pt, kfs] : wide_kfs) n_wide += kfs.size() * (kfs.size() + 1) / 2;
    CHECK(wide_count.size() == n_wide, "gather plan: blocks of a wide point outside its key frames");
    g_wide_points = std::max(g_wide_points, static_cast<int>(wide_kfs.size()));
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
      {"A", b.A, 8 * a}, {"gfirst", b.gfirst, 4 * gf}, {"gcon", b.gcon, 16 * gc}, {"chunk_map", b.chunk_map, 4 * cmap}};
  // the pair streams alias the observation streams when every pair has one observation
  if (b.pair_xy == b.obs_xy) {
    CHECK(npairs == nobs, "pair streams alias observation streams of another length");
    rs.erase(std::remove_if(rs.begin(), rs.end(), [](const Range& r) { return !std::strcmp(r.name, "pair_xy") || !std::strcmp(r.name, "pair_info"); }), rs.end());
  }
  check_ranges(rs);
}

// The packed streams of the last batch whose solve was enqueued (copied at its first launch).
static std::vector<std::vector<char>> g_packed;
static std::vector<float4> g_packed_pts;  // the packed points of the last batch
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
  g_packed_pts.assign(b.pts_init, b.pts_init + b.n_points);
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
// The decision also writes the solution streams: here the initial state, so that a read returns the input.
hipError_t launch_decide(const Batch& b, bool, hipStream_t) {
  ++g_launches;
  check_batch(b);
  if (b.out_pts && b.pts_init) std::memcpy(b.out_pts, b.pts_init, sizeof(float4) * b.n_points);
  if (b.out_poses && b.poses_init) std::memcpy(b.out_poses, b.poses_init, sizeof(float4) * 3 * b.n_poses);
  return hipSuccess;
}
hipError_t launch_gather(const Batch& b, hipStream_t) { ++g_launches; check_batch(b); return hipSuccess; }
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
hipError_t launch_adj_linv(const double*, int, double*, hipStream_t) { return hipSuccess; }
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

// ---- stand-ins of the inertial launches: bounds and index checks of the window they are handed ----------------
namespace sba_imu_dev {
static int g_step = 0;
static double g_cost = 0;

static void check_window(const Window& w) {
  const size_t N = w.N, P = w.P, K = w.K, n = w.n, np = w.npairs, n15 = 15 * n;
  auto atl = [](size_t v) { return std::max<size_t>(v, 1); };
  if (!w.pose_obs || !inside(w.pose_obs, 4 * (n + 1)) || !w.blk_tri || !inside(w.blk_tri, 4 * (w.nblocks + 1))) {
    CHECK(false, "inertial window index heads not bound");
    return;
  }
  const size_t nkobs = atl(w.pose_obs[n]), ntri = atl(w.blk_tri[w.nblocks]);
  check_ranges({{"xy", w.xy, 16 * N}, {"info", w.info, 32 * N}, {"opid", w.opid, 4 * N}, {"okid", w.okid, 4 * N},
                {"ocid", w.ocid, 4 * N}, {"cam", w.cam, 96 * size_t(w.C)}, {"point_obs", w.point_obs, 4 * (P + 1)},
                {"pobs", w.pobs, 4 * N}, {"point_pair", w.point_pair, 4 * (P + 1)},
                {"pair_obs", w.pair_obs, 8 * atl(np)}, {"pair_pose", w.pair_pose, 4 * atl(np)},
                {"pair_point", w.pair_point, 4 * atl(np)}, {"pose_obs", w.pose_obs, 4 * (n + 1)},
                {"kobs", w.kobs, 4 * nkobs}, {"pose_pair", w.pose_pair, 4 * (n + 1)}, {"kpair", w.kpair, 4 * atl(np)},
                {"blk_jk", w.blk_jk, 8 * atl(w.nblocks)}, {"blk_tri", w.blk_tri, 4 * (w.nblocks + 1)},
                {"tri", w.tri, 8 * ntri}, {"pre", w.pre, sizeof(Preint) * K}, {"pts0", w.pts[0], 24 * P},
                {"pts1", w.pts[1], 24 * P}, {"pose0", w.pose[0], 8 * kPS * K}, {"pose1", w.pose[1], 8 * kPS * K},
                {"om", w.om, sizeof(ObsModel) * N}, {"em", w.em, sizeof(EdgeModel) * K}, {"U", w.U, 8 * n15 * n15},
                {"ea", w.ea, 8 * n15}, {"V", w.V, 72 * P}, {"eb", w.eb, 24 * P}, {"W", w.W, 144 * atl(np)},
                {"Vinv", w.Vinv, 72 * P}, {"rp", w.rp, 24 * P}, {"Y", w.Y, 144 * atl(np)},
                {"A", w.A, 8 * 256 * atl(w.nblocks)},
                {"rhs", w.rhs, 8 * n15}, {"dp", w.dp, 24 * P}, {"ocost", w.ocost, 8 * N}, {"ecost", w.ecost, 8 * K},
                {"ppred", w.ppred, 8 * P}, {"pmax", w.pmax, 8 * P}, {"cpred", w.cpred, 8 * n15},
                {"out", w.out, sizeof(IterScalars)}, {"chol_ok", w.chol_ok, 4}});
  CHECK(n <= size_t(kMaxFree) && w.F + w.n == w.K, "free key frames");
  for (size_t o = 0; o < N; ++o)
    CHECK(w.opid[o] >= 0 && size_t(w.opid[o]) < P && w.okid[o] >= 0 && size_t(w.okid[o]) < K && w.ocid[o] >= 0 &&
              w.ocid[o] < w.C, "observation indexes");
  CHECK(w.point_obs[0] == 0 && size_t(w.point_obs[P]) == N, "point ranges");
  std::vector<int> seen(N, 0);
  for (size_t i = 0; i < P; ++i) {
    CHECK(w.point_obs[i] <= w.point_obs[i + 1], "point ranges increase");
    for (int q = w.point_obs[i]; q < w.point_obs[i + 1]; ++q) {
      CHECK(size_t(w.opid[w.pobs[q]]) == i, "pobs groups observations by point");
      ++seen[w.pobs[q]];
    }
    for (int p = w.point_pair[i]; p < w.point_pair[i + 1]; ++p) {
      CHECK(size_t(w.pair_point[p]) == i && w.pair_pose[p] >= 0 && size_t(w.pair_pose[p]) < n, "pair indexes");
      CHECK(w.pair_obs[2 * p] >= w.point_obs[i] && w.pair_obs[2 * p] < w.pair_obs[2 * p + 1] &&
                w.pair_obs[2 * p + 1] <= w.point_obs[i + 1], "pair observation range inside its point's");
      for (int q = w.pair_obs[2 * p]; q < w.pair_obs[2 * p + 1]; ++q)
        CHECK(w.okid[w.pobs[q]] == w.F + w.pair_pose[p], "pair observations of its key frame");
    }
  }
  for (size_t o = 0; o < N; ++o) CHECK(seen[o] == 1, "pobs is a permutation");
  CHECK(size_t(w.point_pair[P]) == np && size_t(w.pose_pair[n]) == np, "pair counts");
  for (size_t j = 0; j < n; ++j) {
    for (int q = w.pose_obs[j]; q < w.pose_obs[j + 1]; ++q) CHECK(size_t(w.okid[w.kobs[q]]) == w.F + j, "kobs");
    for (int q = w.pose_pair[j]; q < w.pose_pair[j + 1]; ++q) CHECK(size_t(w.pair_pose[w.kpair[q]]) == j, "kpair");
  }
  CHECK(size_t(w.nblocks) == n * (n + 1) / 2, "block count");
  for (int b = 0; b < w.nblocks; ++b) {
    const int j = w.blk_jk[2 * b], k = w.blk_jk[2 * b + 1];
    CHECK(j >= k && k >= 0 && size_t(j) < n, "block (j, k)");
    for (int t = w.blk_tri[b]; t < w.blk_tri[b + 1]; ++t) {
      const int2 pr = w.tri[t];
      CHECK(pr.x >= 0 && size_t(pr.x) < np && pr.y >= 0 && size_t(pr.y) < np, "tri pair indexes");
      CHECK(w.pair_pose[pr.x] == j && w.pair_pose[pr.y] == k && w.pair_point[pr.x] == w.pair_point[pr.y],
            "tri pairs of the block's key frames and one point");
    }
  }
}

hipError_t imu_model(const Window& w, hipStream_t) {
  check_window(w);
  ++sba_dev::g_launches;
  return hipSuccess;
}
hipError_t imu_cost(const Window& w, hipStream_t) {
  check_window(w);
  g_cost = 100.0;
  *w.out = IterScalars{g_cost, 0.0, 0.0, 0.0, 0, 1, {0, 0}};
  return hipSuccess;
}
// every fourth step fails to factor; the others halve the cost with the prediction that makes rho = 1
hipError_t imu_step(const Window& w, hipStream_t) {
  check_window(w);
  const int ok = (++g_step % 4) != 0;
  *w.out = IterScalars{0.5 * g_cost, 0.5 * g_cost, 1.0, 1.0, 0, ok, {0, 0}};
  return hipSuccess;
}
// the accepted state moves the free key frames' gyroscope biases past the re-integration threshold
hipError_t imu_accept(const Window& w, hipStream_t) {
  check_window(w);
  g_cost *= 0.5;
  for (int k = w.F; k < w.K; ++k) w.pose[0][kPS * k + kBg] += 0.01;
  return hipSuccess;
}
}  // namespace sba_imu_dev

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

  g_case = "batch: mapped window with points seen from more than 40 free key frames";
  std::printf("%s\n", g_case.c_str());
  {
    Scene loop;
    make_loop(loop, 48, 3000);
    // the loop window plus 5 points (before the fixed ones) seen from 41 .. 47 key frames (free and fixed)
    const sba_problem_view& v = loop.v;
    const int np0 = v.num_points, nfp = v.num_fixed_points, add = 5, nc = v.num_poses;
    std::vector<float> pts(v.points, v.points + 3 * np0), xy(v.observation_xys, v.observation_xys + 2 * v.num_observations),
        info(v.observation_infos, v.observation_infos + 4 * v.num_observations);
    std::vector<int32_t> pid(v.point_ids, v.point_ids + v.num_observations), qid(v.pose_ids, v.pose_ids + v.num_observations);
    std::vector<int8_t> cid(v.camera_ids, v.camera_ids + v.num_observations);
    for (int& i : pid)
      if (i >= np0 - nfp) i += add;  // the fixed points move behind the new ones
    pts.insert(pts.begin() + 3 * (np0 - nfp), 3 * add, 1.0f);
    for (int k = 0; k < add; ++k) {
      const int i = np0 - nfp + k;
      for (int j = 0; j < std::min(nc, 41 + k); ++j) {
        const int jj = (j * 7 + k) % nc;  // key frames in scattered order
        pid.push_back(i);
        qid.push_back(jj);
        cid.push_back(static_cast<int8_t>(j % v.num_cameras));
        xy.push_back(0.01f * j);
        xy.push_back(-0.01f * k);
        for (float x : {1e4f, 0.f, 0.f, 1e4f}) info.push_back(x);
      }
    }
    sba_problem_view w = v;
    w.num_points = np0 + add;
    w.points = pts.data();
    w.num_observations = static_cast<int>(pid.size());
    w.observation_xys = xy.data();
    w.observation_infos = info.data();
    w.point_ids = pid.data();
    w.pose_ids = qid.data();
    w.camera_ids = cid.data();
    Scene small;
    make_scene(small, 8, 500, 41);
    sba_problem_view vs[2] = {small.v, w};
    g_wide_points = 0;
    std::vector<std::vector<char>> packed[2];
    for (int threads : {1, 8}) {
      sba_options o1 = o;
      o1.pack_threads = threads;
      sba_handle* h1 = nullptr;
      CHECK(sba_create(&o1, &h1) == SBA_OK, "sba_create");
      sba_batch* b = nullptr;
      CHECK(sba_batch_create(h1, vs, 2, &b) == SBA_OK, sba_last_error(h1));
      CHECK(sba_batch_run(h1, b) == SBA_OK, "run");
      packed[threads == 8] = g_packed;
      sba_batch_destroy(h1, b);
      std::vector<float> rp(3 * w.num_points), rq(16 * w.num_poses);
      sba_result r{};
      r.points = rp.data();
      r.rig_from_world = rq.data();
      CHECK(sba_solve(h1, &w, &r) == SBA_OK, sba_last_error(h1));
      sba_destroy(h1);
    }
    int expect = 0;  // points seen from more than 40 free key frames
    for (int k = 0; k < add; ++k) {
      int f = 0;
      for (int j = 0; j < std::min(nc, 41 + k); ++j) f += (j * 7 + k) % nc < nc - v.num_fixed_key_frames;
      expect += f > 40;
    }
    // the mapped window's points are packed in another order: a read (of the unchanged state) returns them in
    // the caller's order
    {
      sba_batch* b = nullptr;
      CHECK(sba_batch_create(h, &w, 1, &b) == SBA_OK, sba_last_error(h));
      CHECK(sba_batch_run(h, b) == SBA_OK, "run");
      std::vector<float> rp(3 * w.num_points, -1.f), rq(16 * w.num_poses);
      sba_result r{};
      r.points = rp.data();
      r.rig_from_world = rq.data();
      CHECK(sba_batch_read(h, b, &r) == SBA_OK, "read");
      CHECK(std::memcmp(rp.data(), pts.data(), sizeof(float) * rp.size()) == 0, "points come back in the caller's order");
      bool moved = false;  // the packed order differs from the caller's
      for (int i = 0; i < w.num_points && !moved; ++i)
        moved = std::memcmp(&g_packed_pts[i], &pts[3 * i], sizeof(float) * 3) != 0;
      CHECK(moved, "the mapped window's points are packed in another order");
      sba_batch_destroy(h, b);
    }
    std::printf("  wide points %d (expected %d)\n", g_wide_points, expect);
    CHECK(expect >= 3 && g_wide_points == expect, "the wide points have tiles of their own");
    CHECK(packed[0] == packed[1], "packing with wide points on 1 and on 8 host threads");
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

  g_case = "a batch packed window by window on 1 and on 8 host threads gives the same streams";
  std::printf("%s\n", g_case.c_str());
  {
    std::vector<Scene> scs(7);
    std::vector<sba_problem_view> vs;
    for (int k = 0; k < 6; ++k) {
      make_scene(scs[k], 5 + 6 * k, 300 + 700 * k, 60 + k);
      vs.push_back(scs[k].v);
    }
    make_loop(scs[6], 48, 3000);  // a mapped window among them
    vs.push_back(scs[6].v);
    std::vector<std::vector<char>> packed[2];
    for (int threads : {1, 8}) {
      sba_options o1 = o;
      o1.pack_threads = threads;
      sba_handle* h1 = nullptr;
      CHECK(sba_create(&o1, &h1) == SBA_OK, "sba_create");
      for (int rep = 0; rep < 2; ++rep) {  // the second batch reuses the handle's packing buffers
        sba_batch* b = nullptr;
        CHECK(sba_batch_create(h1, vs.data(), static_cast<int>(vs.size()), &b) == SBA_OK, sba_last_error(h1));
        if (rep == 1) packed[threads == 8] = g_packed;
        sba_batch_destroy(h1, b);
      }
      sba_destroy(h1);
    }
    CHECK(packed[0].size() == packed[1].size() && !packed[0].empty(), "stream count");
    for (size_t k = 0; k < packed[0].size() && k < packed[1].size(); ++k)
      CHECK(packed[0][k] == packed[1][k], (std::string("stream ") + std::to_string(k) + " differs").c_str());
  }

  g_case = "invalid windows are rejected before any launch";
  std::printf("%s\n", g_case.c_str());
  {
    const int launches = g_launches;
    for (int which : {0, 1, 2, 3}) {
      Scene sc;
      make_scene(sc, 8, 400, 70 + which);
      std::vector<int32_t> pid(sc.v.point_ids, sc.v.point_ids + sc.v.num_observations);
      std::vector<int32_t> qid(sc.v.pose_ids, sc.v.pose_ids + sc.v.num_observations);
      std::vector<int8_t> cid(sc.v.camera_ids, sc.v.camera_ids + sc.v.num_observations);
      const int o = sc.v.num_observations / 2;
      if (which == 0) pid[o] = sc.v.num_points;
      if (which == 1) qid[o] = -1;
      if (which == 2) cid[o] = static_cast<int8_t>(sc.v.num_cameras);
      sba_problem_view bad = sc.v;
      bad.point_ids = pid.data();
      bad.pose_ids = qid.data();
      bad.camera_ids = cid.data();
      if (which == 3) bad.num_fixed_key_frames = bad.num_poses + 1;
      sba_result r = result_for(sc);
      CHECK(sba_solve(h, &bad, &r) == SBA_E_INVALID, "single window");
      Scene good;
      make_scene(good, 6, 300, 80 + which);
      sba_problem_view vs[3] = {good.v, bad, good.v};
      sba_batch* b = nullptr;
      CHECK(sba_batch_create(h, vs, 3, &b) == SBA_E_INVALID && b == nullptr, "window 1 of a batch");
      CHECK(std::string(sba_last_error(h)).find("window 1") != std::string::npos || which == 2,
            "the error names the window");
    }
    CHECK(g_launches == launches, "no launch for a rejected window");
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
  g_case = "inertial solve: packing, layout, re-integration";
  std::printf("%s\n", g_case.c_str());
  {
    const int K = 6, F = 2, P = 40, S = 40;
    sba_imu_calibration cal{};
    for (int i = 0; i < 4; ++i) cal.rig_from_imu[5 * i] = 1.f;
    cal.gyroscope_noise_density = 1.7e-4f;
    cal.gyroscope_random_walk = 1.9e-5f;
    cal.accelerometer_noise_density = 2e-3f;
    cal.accelerometer_random_walk = 3e-3f;
    cal.frequency = 200.f;
    std::vector<sba_imu_measurement> meas(S * (K - 1));
    for (size_t i = 0; i < meas.size(); ++i) {
      meas[i].time_ns = int64_t(i + 1) * 5000000;
      meas[i].linear_acceleration[1] = -9.81f;
      meas[i].angular_velocity[0] = 0.1f;
    }
    std::vector<sba_imu_preintegration> pre(K);
    const float zero[3] = {0, 0, 0};
    for (int k = 0; k < K; ++k) {
      const int n = k + 1 < K ? S : 0;
      CHECK(sba_imu_integrate(&cal, meas.data() + S * k, n, zero, zero, &pre[k]) == SBA_OK, "integrate");
      pre[k].first_measurement = S * k;
    }
    std::vector<float> pts(3 * P), w_from_imu(16 * K, 0.f), vel(3 * K, 0.f), bg(3 * K, 0.f), ba(3 * K, 0.f);
    for (int i = 0; i < P; ++i) {
      pts[3 * i] = 0.1f * (i % 7) - 0.3f;
      pts[3 * i + 1] = 0.1f * (i % 5) - 0.2f;
      pts[3 * i + 2] = 4.f + 0.05f * i;
    }
    for (int k = 0; k < K; ++k) {
      for (int i = 0; i < 4; ++i) w_from_imu[16 * k + 5 * i] = 1.f;
      w_from_imu[16 * k + 12] = 0.1f * k;
    }
    std::vector<float> xy, info;
    std::vector<int32_t> pid, kid;
    std::vector<int8_t> cid;
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < P; ++i) {
        if ((i + k) % 3 == 0) continue;  // points seen by some key frames only
        xy.push_back(0.f);
        xy.push_back(0.f);
        for (float f : {1e4f, 0.f, 0.f, 1e4f}) info.push_back(f);
        pid.push_back(i);
        kid.push_back(k);
        cid.push_back(int8_t(i % 2));
      }
    std::vector<float> cams(32, 0.f);
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < 4; ++i) cams[16 * c + 5 * i] = 1.f;
    cams[16 + 12] = -0.12f;
    sba_imu_problem_view v{};
    v.num_points = P;
    v.num_observations = int(pid.size());
    v.num_poses = K;
    v.num_fixed_key_frames = F;
    v.num_cameras = 2;
    v.num_measurements = int(meas.size());
    v.points = pts.data();
    v.observation_xys = xy.data();
    v.observation_infos = info.data();
    v.point_ids = pid.data();
    v.pose_ids = kid.data();
    v.camera_ids = cid.data();
    v.camera_from_rig = cams.data();
    v.w_from_imu = w_from_imu.data();
    v.velocity = vel.data();
    v.gyro_bias = bg.data();
    v.acc_bias = ba.data();
    v.preintegration = pre.data();
    v.measurements = meas.data();
    v.gravity[1] = 9.81f;
    v.max_iterations = 7;
    v.robustifier_scale = 1.5f;
    v.imu_penalty = 1e-2f;
    v.acc_rw_penalty = 0.1f;
    v.boundary_imu_penalty = 1e-2f;
    v.reintegration_thresh = 1e-3f;
    std::vector<float> opts(3 * P), opose(16 * K), ovel(3 * K), obg(3 * K), oba(3 * K);
    std::vector<sba_imu_preintegration> opre(K);
    sba_imu_result r{};
    r.points = opts.data();
    r.w_from_imu = opose.data();
    r.velocity = ovel.data();
    r.gyro_bias = obg.data();
    r.acc_bias = oba.data();
    r.preintegration = opre.data();
    sba_imu_dev::g_step = 0;
    CHECK(sba_imu_solve(h, &cal, &v, &r) == SBA_OK, sba_last_error(h));
    CHECK(r.iterations == 7 && r.accepted == 6, "iterations: 7, the 4th failing to factor");
    for (int k = F; k < K - 1; ++k)  // the biases moved by 0.01 per accepted step: re-integrated there
      CHECK(std::fabs(opre[k].gyro_bias[0] - obg[3 * k]) < 1e-6f && std::fabs(obg[3 * k] - 0.06f) < 1e-6f,
            "preintegration re-integrated at the final biases");
    CHECK(opre[0].gyro_bias[0] == 0.f, "fixed key frames keep their preintegration");
    v.num_fixed_key_frames = 0;
    v.num_poses = K;
    CHECK(sba_imu_solve(h, &cal, &v, &r) == SBA_OK, sba_last_error(h));
    v.num_fixed_key_frames = F;
    v.pose_ids = nullptr;
    CHECK(sba_imu_solve(h, &cal, &v, &r) == SBA_E_INVALID, "null buffer");
  }

  sba_destroy(h);
  g_case = "leaks";
  // the fake scratch stays (it lives as long as the process); everything else is freed
  CHECK(g_allocs.size() == 1, "every allocation freed");
  std::printf("%s (%d failed checks, %d launches)\n", g_fail ? "FAILED" : "OK", g_fail, sba_dev::g_launches);
  return g_fail ? 1 : 0;
}
