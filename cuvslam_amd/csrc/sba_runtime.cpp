// Host runtime of the MI355X SBA: the C ABI of include/cuvslam_sba.h.
//
// sba_batch_create packs windows into the structure-of-arrays streams of sba_device.h (a counting sort of
// each window's observations by point, then by key frame and camera -- the reference does the same sort on
// the host, libs/cuda_modules/sba.cpp:451-535), plans tiles and chunks, uploads once, and records the
// whole solve (initial cost, max_iterations LM iterations, final gather) as one hipGraph on the handle's
// stream. sba_batch_run replays it; nothing returns to the host between iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/cuvslam_sba.h"
#include "sba_device.h"

namespace sba_dev {
hipError_t launch_reset(const Batch& b, hipStream_t s);
hipError_t launch_schur(const Batch& b, hipStream_t s);
hipError_t launch_reduce(const Batch& b, hipStream_t s);
hipError_t launch_factor(const Batch& b, double* dbg_S, hipStream_t s, double* store_L = nullptr);
hipError_t launch_adj_obs(const Batch& b, int mode, const double* spts, const double* sposes, const double* v, double* y,
                          double* ec, const double* zc, double* dxy, double* dinfo, const int32_t* orig, int32_t* done,
                          hipStream_t s);
hipError_t launch_adj_state(const Batch& b, double* spts, double* sposes, const double* a, int op, hipStream_t s);
hipError_t launch_adj_linv(const double* Lst, int n, double* Linv, hipStream_t s);
hipError_t launch_adj_chol(const double* Linv, int n, const double* ec, double* zc, const int32_t* done, hipStream_t s);
hipError_t launch_adj_vec(const Batch& b, double* x, double* r, double* z, double* p, double* y, double* ec, double* zc,
                          double* acc, double* g, double* sc, int32_t* done, double tol, int nx, int n6, int op,
                          hipStream_t s);
hipError_t launch_backsub(const Batch& b, bool initial, hipStream_t s);
hipError_t launch_large_factor(const Batch& b, int w, int n, hipStream_t s);
hipError_t launch_winsorize_select(const Batch& b, int w, int np, int kf, double* err, int32_t* pt_of,
                                   int32_t* count, uint8_t* mark, int32_t* out, hipStream_t s);
hipError_t launch_winsorize_apply(const Batch& b, int w, int np, int nc, const uint8_t* mark, hipStream_t s);
hipError_t launch_decide(const Batch& b, bool initial, hipStream_t s);
hipError_t launch_gather(const Batch& b, hipStream_t s);
hipError_t launch_copy_calibrate(float4* dst, const float4* src, int64_t n, hipStream_t s);
hipError_t launch_mfma_layout(const float* A, const float* B, float* D, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);
int scratch_setup(int device, void** ptr, size_t* bytes);
hipError_t set_kernel_attributes();
}  // namespace sba_dev

using namespace sba_dev;

// Page-locked host memory for the streams that are uploaded: their copies to the device are then DMA transfers at
// the link's rate instead of staged copies of pageable memory.
template <class T>
struct PinnedAllocator {
  using value_type = T;
  PinnedAllocator() = default;
  template <class U>
  PinnedAllocator(const PinnedAllocator<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess || !p)
      throw std::bad_alloc();
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  // resize() default-initializes: the packing writes every element, so the streams are not cleared first
  template <class U>
  void construct(U* p) {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
  template <class U>
  bool operator==(const PinnedAllocator<U>&) const { return true; }
  template <class U>
  bool operator!=(const PinnedAllocator<U>&) const { return false; }
};
template <class T>
using pinned_vector = std::vector<T, PinnedAllocator<T>>;

// Host threads kept by a handle: run(n, fn) calls fn(0) .. fn(n - 1), fn(0) on the calling thread, and returns when
// all have returned. Threads are started on first need and end with the handle, so a solve does not pay for
// thread creation. Between jobs a thread polls for the next one for up to kSpin before it sleeps: the packing
// issues its phases back to back, and waking sleeping threads for each would cost more than the phases. Not
// reentrant: fn must not call run().
class HostPool {
 public:
  HostPool() = default;
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_.store(true);
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
  }
  void run(int n, const std::function<void(int)>& fn) {
    if (n <= 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      while (static_cast<int>(threads_.size()) < n - 1) {
        const int id = static_cast<int>(threads_.size()) + 1;
        const uint64_t seen = gen_.load();
        threads_.emplace_back([this, id, seen] { worker(id, seen); });
      }
      job_ = &fn;
      job_n_ = n;
      pending_.store(static_cast<int>(threads_.size()));
      gen_.fetch_add(1, std::memory_order_release);
    }
    wake_.notify_all();
    fn(0);
    const auto t0 = std::chrono::steady_clock::now();
    while (pending_.load(std::memory_order_acquire) != 0) {
      if (std::chrono::steady_clock::now() - t0 > kSpin) {
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return pending_.load(std::memory_order_acquire) == 0; });
        break;
      }
      std::this_thread::yield();
    }
  }

 private:
  static constexpr std::chrono::microseconds kSpin{2000};
  void worker(int id, uint64_t seen) {
    for (;;) {
      const auto t0 = std::chrono::steady_clock::now();
      while (gen_.load(std::memory_order_acquire) == seen && !stop_.load()) {
        if (std::chrono::steady_clock::now() - t0 > kSpin) {
          std::unique_lock<std::mutex> lk(m_);
          wake_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
          break;
        }
        std::this_thread::yield();
      }
      if (stop_.load()) return;
      seen = gen_.load(std::memory_order_acquire);
      if (id < job_n_) (*job_)(id);
      if (pending_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        std::lock_guard<std::mutex> g(m_);
        done_.notify_one();
      }
    }
  }
  std::mutex m_;
  std::condition_variable wake_, done_;
  std::vector<std::thread> threads_;
  const std::function<void(int)>* job_ = nullptr;
  int job_n_ = 0;
  std::atomic<int> pending_{0};
  std::atomic<uint64_t> gen_{0};
  std::atomic<bool> stop_{false};
};

// One window's sort order and pair counts, from the first packing phase to the third.
struct WindowSort {
  std::vector<int32_t> cnt, cur, start, npair, nfree, first_pair;
  std::vector<int64_t> ko;    // sorted (key << 32 | input index) of the observations
  std::vector<int64_t> priv;  // the threads' ranges, each sorted by point
  int npairs = 0;
  // mapped windows: the points seen from more than kMaxFreePoses free key frames ("wide" points, staged by tiles
  // of their own), the pair copies their tiles need and where those start in the pair streams
  std::vector<int32_t> wide;
  int64_t ndup = 0, dup_off = 0;
  // mapped windows: packed point -> window point, the free points before the fixed ones, each class ordered by
  // its first key frame (stable), so that a chunk's consecutive points share key frames; empty otherwise (the
  // packed order is the window's). npair, nfree, first_pair and wide follow the packed order, start the window's.
  std::vector<int32_t> order;
};

// A wide point's free key frames are split into groups of at most kWideGroup consecutive ones; every two groups
// are staged together by one chunk (at most kMaxFreePoses key frames).
constexpr int kWideGroup = kMaxFreePoses / 2;
int wide_groups(int nfree) { return (nfree + kWideGroup - 1) / kWideGroup; }
// pair_pose of a wide point's pair copy whose free key frame its tile does not stage
constexpr int32_t kNotStaged = INT32_MAX;

// The host side of a batch while it is packed. A handle keeps one and clears it per batch, so that its vectors keep
// their capacity across calls: a repeated solve writes into memory that is already mapped (first-touch page faults
// of fresh buffers cost as much as the packing itself). The uploaded streams are page-locked.
struct HostPack {
  pinned_vector<float2> obs_xy;
  pinned_vector<float4> obs_info;
  pinned_vector<int32_t> pair_obs, pair_pose, point_pair, point_nfree, obs_orig;
  pinned_vector<int4> pair_rec;
  pinned_vector<float2> pair_xy;
  pinned_vector<float4> pair_info;
  pinned_vector<float4> pts, poses, cams;
  pinned_vector<Tile> tiles, btiles;
  pinned_vector<Chunk> chunks, bchunks;
  pinned_vector<int32_t> chunk_map, red_first, gfirst;
  pinned_vector<int4> gcon;
  // per-window scratch of the packing (the sort order and pair counts, the tiles before they are joined)
  std::vector<WindowSort> sorts;
  std::vector<std::vector<Tile>> wtiles, wbtiles;
  void clear() {
    for (auto* v : {&pair_obs, &pair_pose, &point_pair, &point_nfree, &obs_orig}) v->clear();
    obs_xy.clear();
    obs_info.clear();
    pair_rec.clear();
    pair_xy.clear();
    pair_info.clear();
    pts.clear();
    poses.clear();
    cams.clear();
    tiles.clear();
    btiles.clear();
    chunks.clear();
    bchunks.clear();
    chunk_map.clear();
    red_first.clear();
    gfirst.clear();
    gcon.clear();
  }
};

struct sba_handle {
  sba_options opt{};
  hipStream_t stream = nullptr;
  std::string err;
  bool attrs_set = false;
  // device workspace of sba_solve, kept across calls and regrown to twice the need (the reference bundler keeps
  // its buffers the same way, schur_complement_bundler_gpu.cpp:92-99)
  char* ws = nullptr;
  size_t ws_bytes = 0;
  HostPack pack;  // host packing buffers, kept across calls
  HostPool pool;  // host threads of the packing
  // this device's address of the fixed scratch array of sba_solve_noalloc (looked up by sba_create)
  char* scratch = nullptr;
  size_t scratch_bytes = 0;
  // device workspace of the inertial solve (sba_imu_solve), kept and regrown like ws
  char* imu_ws = nullptr;
  size_t imu_ws_bytes = 0;
};

struct sba_batch {
  int n = 0;
  int max_iter = 0;
  std::vector<WindowDesc> win;
  std::vector<int32_t> status_pre;  // host-detected per-window status (empty problems)
  std::vector<void*> allocs;
  Batch dev{};
  int64_t total_obs = 0, total_pairs = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double* dbg_S = nullptr;  // debug buffer of the reduced system of window 0
  int32_t* obs_orig = nullptr;  // packed observation -> index in its window's input order
  std::vector<int> mapped;      // windows with more than kMaxFreePoses free key frames
  std::vector<std::vector<int32_t>> point_order;  // per window: packed point -> window point (empty: the same)
  double* adj = nullptr;        // adjoint workspace of window 0 (allocated on first use)
  bool adj_ready = false;       // the workspace holds the factor and the FP64 state of the current solution
  hipEvent_t ev[8] = {};        // kernel-class timing (created on first use)
  float class_ms[5] = {0, 0, 0, 0, 0};
  int class_launches[5] = {0, 0, 0, 0, 0};
};

namespace {

// A roctx range over a host phase of the solver (the reference traces the same phases,
// schur_complement_bundler_gpu.cpp:116,176); free when no profiler is attached.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

int fail(sba_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return fail(h, SBA_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Runs fn(t, begin, end) over [0, n) cut into `nthreads` contiguous ranges, range t on the handle's pool thread t
// (range 0 on the caller's).
template <class F>
void parallel_ranges(sba_handle* h, int nthreads, int64_t n, const F& fn) {
  if (nthreads <= 1 || n < 2 * nthreads) {
    fn(0, int64_t(0), n);
    return;
  }
  h->pool.run(nthreads, [&](int t) { fn(t, n * t / nthreads, n * (t + 1) / nthreads); });
}

// Host threads of the packing: the option when set, else up to 16 (for one window, only when it has 65536
// observations or more).
int pack_threads(const sba_handle* h, int64_t work) {
  if (h->opt.pack_threads > 0) return h->opt.pack_threads;
  if (work < (1 << 16)) return 1;
  return static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
}

// Packing phase 1 of window w: validates it, fills the sizes of W and sorts its observations into the order
// (point, key frame, camera, input index): a stable counting sort by point (per-thread histograms, then each
// thread scatters its own input range), then within each point a stable sort by (key frame, camera); then the
// pairs (one per run of equal key frames in a point's segment) are counted per point. T threads share the work;
// every output element has one writer, so the result does not depend on T. On failure `err` says why.
int sort_window(sba_handle* h, const sba_problem_view& p, int w, int T, WindowSort& S, WindowDesc& W, std::string& err) {
  std::unique_ptr<TraceRange> phase;
#define PHASE(name) phase.reset(), phase.reset(new TraceRange("sba sort: " name))
  const int np = p.num_points, nc = p.num_poses, no = p.num_observations, ncam = p.num_cameras;
  const std::string where = " in window " + std::to_string(w);
  if (np < 0 || nc < 0 || no < 0 || ncam < 0) return err = "negative size" + where, SBA_E_INVALID;
  if (p.num_fixed_points < 0 || p.num_fixed_points > np || p.num_fixed_key_frames < 0 || p.num_fixed_key_frames > nc)
    return err = "fixed counts out of range" + where, SBA_E_INVALID;
  if ((np && !p.points) || (nc && !p.rig_from_world) || (ncam && !p.camera_from_rig) ||
      (no && (!p.observation_xys || !p.observation_infos || !p.point_ids || !p.pose_ids || !p.camera_ids)))
    return err = "null buffer" + where, SBA_E_INVALID;
  const int ncf = nc - p.num_fixed_key_frames;
  if (nc > kMaxPoses)
    return err = "window " + std::to_string(w) + " has more than " + std::to_string(kMaxPoses) + " key frames", SBA_E_CAPACITY;
  if (ncam > kMaxCameras) return err = "more than 32 cameras", SBA_E_CAPACITY;
  W = WindowDesc{};
  W.np = np;
  W.npf = np - p.num_fixed_points;
  W.nc = nc;
  W.ncf = ncf;
  W.ncam = ncam;
  W.no = no;
  W.max_iterations = p.max_iterations;
  W.robust = p.robustifier_scale;
  W.mapped = ncf > kMaxFreePoses ? 1 : 0;

  PHASE("histogram");
  // The sort by point: each thread counts, then sorts its own contiguous range of the input into its own range of
  // a private buffer (a counting sort whose writes stay in the thread's range: threads never write the same cache
  // lines), and the point segments are then gathered from the threads' ranges in thread order, i.e. in input order.
  // Each word carries the observation's sort key next to its index: key (key frame << 8 | camera) above, input
  // index below, so that the sort within a point is a sort of the words (stable by key) and the later passes read
  // the keys in order instead of the input at random.
  const int TT = (T <= 1 || no < 2 * T) ? 1 : T;  // the ranges parallel_ranges makes
  auto range_begin = [&](int t) { return static_cast<int64_t>(no) * t / TT; };
  const int64_t stride = static_cast<int64_t>(np) + 1;
  std::vector<int32_t>& cnt = S.cnt;  // row t: thread t's counts by point, then its first slot of each point
  cnt.assign(static_cast<size_t>(TT) * stride, 0);
  std::vector<uint8_t> bad(TT, 0);
  auto each_thread = [&](const std::function<void(int)>& fn) { h->pool.run(TT, fn); };
  each_thread([&](int t) {
    int32_t* c = cnt.data() + t * stride;
    for (int64_t o = range_begin(t); o < range_begin(t + 1); ++o) {
      const int32_t i = p.point_ids[o];
      if (i < 0 || i >= np || p.pose_ids[o] < 0 || p.pose_ids[o] >= nc || p.camera_ids[o] < 0 || p.camera_ids[o] >= ncam)
        bad[t] = 1;
      else
        ++c[i];
    }
  });
  if (std::find(bad.begin(), bad.end(), 1) != bad.end()) return err = "observation index out of range" + where, SBA_E_INVALID;
  PHASE("slots");
  std::vector<int32_t>& cur = S.cur;
  cur.resize(static_cast<size_t>(TT) * stride);
  each_thread([&](int t) {  // thread t's first slot of each point in its own range, and a cursor copy
    int32_t* c = cnt.data() + t * stride;
    int32_t run = 0;
    for (int i = 0; i < np; ++i) {
      const int32_t k = c[i];
      c[i] = run;
      run += k;
    }
    c[np] = run;
    std::copy(c, c + stride, cur.data() + t * stride);
  });
  std::vector<int32_t>& start = S.start;
  start.resize(np + 1);
  parallel_ranges(h, T, np, [&](int, int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int32_t s = 0;
      for (int t = 0; t < TT; ++t) s += cnt[t * stride + i + 1] - cnt[t * stride + i];
      start[i] = s;
    }
  });
  {
    int32_t run = 0;
    for (int i = 0; i < np; ++i) {
      const int32_t c = start[i];
      start[i] = run;
      run += c;
    }
    start[np] = run;
  }
  PHASE("scatter");
  std::vector<int64_t>&ko = S.ko, &priv = S.priv;
  ko.resize(no);
  priv.resize(no);
  each_thread([&](int t) {
    int32_t* c = cur.data() + t * stride;
    int64_t* out = priv.data() + range_begin(t);
    for (int64_t o = range_begin(t); o < range_begin(t + 1); ++o)
      out[c[p.point_ids[o]]++] = (static_cast<int64_t>((p.pose_ids[o] << 8) | p.camera_ids[o]) << 32) | o;
  });
  parallel_ranges(h, T, np, [&](int, int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int64_t* dst = ko.data() + start[i];
      for (int t = 0; t < TT; ++t) {
        const int64_t* src = priv.data() + range_begin(t);
        for (int32_t k = cnt[t * stride + i]; k < cnt[t * stride + i + 1]; ++k) *dst++ = src[k];
      }
    }
  });
  PHASE("segments");
  parallel_ranges(h, T, np, [&](int, int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int64_t* seg = ko.data() + start[i];
      const int m = start[i + 1] - start[i];
      if (m <= 48) {  // insertion sort: linear on segments already in key-frame order
        for (int x = 1; x < m; ++x) {
          const int64_t v = seg[x];
          int y = x;
          for (; y > 0 && seg[y - 1] > v; --y) seg[y] = seg[y - 1];
          seg[y] = v;
        }
      } else {
        std::sort(seg, seg + m);
      }
    }
  });
  PHASE("pairs");
  std::vector<int32_t>&npair = S.npair, &nfree = S.nfree;
  npair.resize(np);
  nfree.resize(np);
  parallel_ranges(h, T, np, [&](int, int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int c = 0, f = 0, last = -1;
      for (int k = start[i]; k < start[i + 1]; ++k) {
        const int pose = static_cast<int>(ko[k] >> 40);
        if (pose != last) {
          ++c;
          f += pose < ncf;
          last = pose;
        }
      }
      npair[i] = c;
      nfree[i] = f;
    }
  });
  // mapped windows: the packed point order (a counting sort by first key frame)
  std::vector<int32_t>& order = S.order;
  order.clear();
  if (W.mapped && np > 0) {
    std::vector<int32_t> bucket(2 * static_cast<size_t>(nc) + 1, 0);
    const int npf = W.npf;
    auto key = [&](int o) { return (o < npf ? 0 : nc) + (start[o] < start[o + 1] ? static_cast<int>(ko[start[o]] >> 40) : 0); };
    for (int o = 0; o < np; ++o) ++bucket[key(o) + 1];
    for (size_t k = 1; k < bucket.size(); ++k) bucket[k] += bucket[k - 1];
    order.resize(np);
    for (int o = 0; o < np; ++o) order[bucket[key(o)]++] = o;
    std::vector<int32_t> tmp(np);
    for (int i = 0; i < np; ++i) tmp[i] = npair[order[i]];
    npair.swap(tmp);
    for (int i = 0; i < np; ++i) tmp[i] = nfree[order[i]];
    nfree.swap(tmp);
  }
  // a mapped window's chunks stage at most kMaxFreePoses free key frames: points seen from more are wide
  S.wide.clear();
  S.ndup = 0;
  if (W.mapped)
    for (int i = 0; i < np; ++i)
      if (nfree[i] > kMaxFreePoses) {
        const int64_t K = wide_groups(nfree[i]);
        S.wide.push_back(i);
        S.ndup += K * (K - 1) / 2 * npair[i];
      }
  std::vector<int32_t>& first_pair = S.first_pair;
  first_pair.assign(np + 1, 0);
  for (int i = 0; i < np; ++i) first_pair[i + 1] = first_pair[i] + npair[i];
  S.npairs = first_pair[np];
  return SBA_OK;
}

// Packing phase 3 of window w: writes its streams at the offsets that phase 2 put in W (observations in sorted
// order, pairs, points, key frames, cameras) and builds its Schur and back-substitution tiles (returned in `tiles`
// and `btiles`, with batch-wide indexes) and pair records. Only the window's own ranges of the streams are touched,
// so windows can be written concurrently.
void fill_window(sba_handle* h, const sba_problem_view& p, int w, int T, const WindowSort& S, const WindowDesc& W,
                 HostPack& H, std::vector<Tile>& tiles, std::vector<Tile>& btiles) {
  const int np = W.np, nc = W.nc, no = W.no, ncam = W.ncam, ncf = W.ncf;
  const int obs_base = W.obs_off, pair_base = W.pair_off, point_base = W.point_off;
  const std::vector<int32_t>&start = S.start, &first_pair = S.first_pair;
  const std::vector<int64_t>& ko = S.ko;
  // packed point i is window point orig(i)
  auto orig = [&](int i) { return S.order.empty() ? i : S.order[i]; };
  // the gather into sorted order reads the input at random (disjoint output ranges per thread)
  parallel_ranges(h, T, no, [&](int, int64_t a, int64_t e) {
    for (int64_t k = a; k < e; ++k) {
      const int o = static_cast<int32_t>(ko[k]);
      H.obs_orig[obs_base + k] = o;
      H.obs_xy[obs_base + k] = make_float2(p.observation_xys[2 * o], p.observation_xys[2 * o + 1]);
      const float* L = p.observation_infos + 4 * o;  // column-major (L00, L10, L01, L11)
      float4 inf;
      inf.x = L[0];
      inf.y = 0.5f * (L[1] + L[2]);
      inf.z = L[3];
      int cam = static_cast<int>((ko[k] >> 32) & 255);
      std::memcpy(&inf.w, &cam, 4);
      H.obs_info[obs_base + k] = inf;
    }
  });
  parallel_ranges(h, T, np, [&](int, int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int q = pair_base + first_pair[i];
      H.point_pair[point_base + i] = q;
      H.point_nfree[point_base + i] = S.nfree[i];
      const int o = orig(static_cast<int>(i));
      H.pts[point_base + i] = make_float4(p.points[3 * o], p.points[3 * o + 1], p.points[3 * o + 2], 0.f);
      int last = -1;
      for (int k = start[o]; k < start[o + 1]; ++k) {
        const int pose = static_cast<int>(ko[k] >> 40);
        if (pose != last) {
          H.pair_obs[q] = obs_base + k;
          H.pair_pose[q] = pose;
          ++q;
          last = pose;
        }
      }
    }
  });
  auto put_pose = [](float4* dst, const float* m) {
    for (int r = 0; r < 3; ++r) dst[r] = make_float4(m[r], m[4 + r], m[8 + r], m[12 + r]);
  };
  for (int j = 0; j < nc; ++j) put_pose(H.poses.data() + 3 * (W.pose_off + j), p.rig_from_world + 16 * j);
  for (int c = 0; c < ncam; ++c) put_pose(H.cams.data() + 3 * (W.cam_off + c), p.camera_from_rig + 16 * c);
  // tiles: consecutive points with at most kTilePairs pairs and max_points points; in a mapped window also at
  // most kMaxFreePoses distinct free key frames (one tile is staged by one chunk), and without the wide points
  std::vector<uint8_t> in_tile(W.mapped ? ncf : 0);
  auto wide = [&](int i) { return W.mapped && S.nfree[i] > kMaxFreePoses; };
  auto make_tiles = [&](std::vector<Tile>& out, int max_points) {
    const bool schur = max_points == kTilePoints;
    int i = 0;
    while (i < np) {
      if (schur && wide(i)) {
        ++i;
        continue;
      }
      const int begin = i;
      int pairs = 0, nkf = 0;
      std::fill(in_tile.begin(), in_tile.end(), 0);
      while (i < np && i - begin < max_points && !(schur && wide(i))) {
        const int pi = S.npair[i];
        if (pairs + pi > kTilePairs && i > begin) break;
        if (W.mapped && schur) {
          const int q0 = pair_base + first_pair[i];
          int add = 0;
          for (int q = q0; q < q0 + pi; ++q) {
            const int j = H.pair_pose[q];
            if (j < ncf && !in_tile[j]) ++add;
          }
          if (nkf + add > kMaxFreePoses && i > begin) break;
          for (int q = q0; q < q0 + pi; ++q) {
            const int j = H.pair_pose[q];
            if (j < ncf && !in_tile[j]) in_tile[j] = 1;
          }
          nkf += add;
        }
        pairs += pi;
        ++i;
      }
      const int pb = pair_base + first_pair[begin];
      out.push_back(Tile{w, point_base + begin, point_base + i, pb, pb + pairs, 0, 0, 0});
    }
  };
  tiles.clear();
  btiles.clear();
  make_tiles(tiles, kTilePoints);
  make_tiles(btiles, kBlock);
  // pair records: observation range, key frame | point index inside its back-substitution tile << 16, point
  // index inside its Schur tile | local key frame << 8 (ranges end at the window's own ends)
  const int pair_end = pair_base + S.npairs;
  auto pair_range = [&](int i) {  // batch-wide pair range of batch-wide point i of this window
    const int qa = H.point_pair[i];
    return std::make_pair(qa, i + 1 < point_base + np ? H.point_pair[i + 1] : pair_end);
  };
  // observation end of pair q, the last pair of batch-wide point i ends where the point's observations end
  auto obs_stop = [&](int q, int qe, int i) { return q + 1 < qe ? H.pair_obs[q + 1] : obs_base + start[orig(i - point_base) + 1]; };
  parallel_ranges(h, T, static_cast<int64_t>(tiles.size()), [&](int, int64_t a, int64_t e) {
    for (int64_t t = a; t < e; ++t) {
      const Tile& Tl = tiles[t];
      for (int i = Tl.point_begin; i < Tl.point_end; ++i) {
        const auto [qa, qe] = pair_range(i);
        for (int q = qa; q < qe; ++q) {
          const int ob = H.pair_obs[q];
          const int oe = obs_stop(q, qe, i);
          // local free key frame index: the key frame itself unless the window is mapped (set with its chunks)
          const int jl = H.pair_pose[q] < ncf ? H.pair_pose[q] : 255;
          H.pair_rec[q] = make_int4(ob, oe, H.pair_pose[q], (i - Tl.point_begin) | (jl << 8));
        }
      }
    }
  });
  // the wide points' own pairs (in no Schur tile; the back-substitution reads them)
  for (int i : S.wide) {
    const auto [qa, qe] = pair_range(point_base + i);
    for (int q = qa; q < qe; ++q) {
      const int oe = obs_stop(q, qe, point_base + i);
      H.pair_rec[q] = make_int4(H.pair_obs[q], oe, H.pair_pose[q], 255 << 8);
    }
  }
  parallel_ranges(h, T, static_cast<int64_t>(btiles.size()), [&](int, int64_t a, int64_t e) {
    for (int64_t t = a; t < e; ++t) {
      const Tile& Tl = btiles[t];
      for (int i = Tl.point_begin; i < Tl.point_end; ++i) {
        const auto [qa, qe] = pair_range(i);
        for (int q = qa; q < qe; ++q) H.pair_rec[q].z = H.pair_pose[q] | ((i - Tl.point_begin) << 16);
      }
    }
  });
  // The tiles of the wide points: for groups ga < gb (in this order) a copy of all the point's pairs in which the
  // free key frames outside ga and gb are marked as not staged (pair_pose kNotStaged, local index 255). The
  // first tile that stages a group keeps the blocks inside it (and its key frames' U, e), every tile the blocks
  // between its two groups: each block of the point's Z Z^T is then summed once.
  int64_t d = S.dup_off;
  for (int i : S.wide) {
    const auto [qa, qe] = pair_range(point_base + i);
    const int m = S.nfree[i], K = wide_groups(m);
    // the point's pairs are in key frame order, so its free key frames are its first m pairs
    auto group_begin = [&](int g) { return static_cast<int>(static_cast<int64_t>(g) * m / K); };
    for (int ga = 0; ga < K; ++ga)
      for (int gb = ga + 1; gb < K; ++gb) {
        const int a0 = group_begin(ga), a1 = group_begin(ga + 1), b0 = group_begin(gb), b1 = group_begin(gb + 1);
        for (int q = qa; q < qe; ++q) {
          const int k = q - qa, dq = static_cast<int>(d) + k;
          const bool staged = (k >= a0 && k < a1) || (k >= b0 && k < b1);
          H.pair_obs[dq] = H.pair_obs[q];
          H.pair_pose[dq] = k < m && !staged ? kNotStaged : H.pair_pose[q];
          H.pair_rec[dq] = make_int4(H.pair_rec[q].x, H.pair_rec[q].y, H.pair_rec[q].z, 255 << 8);
        }
        const int virt = 1 | (gb == 1 ? 2 : 0) | (ga == 0 ? 4 : 0) | ((a1 - a0) << 8);
        tiles.push_back(Tile{w, point_base + i, point_base + i + 1, static_cast<int32_t>(d),
                             static_cast<int32_t>(d) + (qe - qa), static_cast<int32_t>(d) - qa, virt, 0});
        d += qe - qa;
      }
  }
}

// Packs the n windows into H (phases 1-3 above, phase 2 the batch-wide offsets). A single window uses the threads
// inside its phases; a batch spreads whole windows over the threads (each window on one thread, taken in turn).
// The packed streams do not depend on the thread counts.
int pack_windows(sba_handle* h, const sba_problem_view* problems, int n, HostPack& H, std::vector<WindowDesc>& win) {
  std::vector<WindowSort>& sorts = H.sorts;
  std::vector<std::vector<Tile>>&tiles = H.wtiles, &btiles = H.wbtiles;
  if (static_cast<int>(sorts.size()) < n) sorts.resize(n);
  if (static_cast<int>(tiles.size()) < n) tiles.resize(n);
  if (static_cast<int>(btiles.size()) < n) btiles.resize(n);
  std::vector<int> rc(n, SBA_OK);
  std::vector<std::string> err(n);
  int64_t total_obs = 0;
  for (int w = 0; w < n; ++w) total_obs += std::max(problems[w].num_observations, 0);
  const int inner = n == 1 ? pack_threads(h, total_obs) : 1;
  const int outer = n == 1 ? 1 : std::min(n, pack_threads(h, total_obs));
  std::atomic<int> next{0};
  auto over_windows = [&](auto&& body) {
    next = 0;
    h->pool.run(outer, [&](int) {
      for (int w = next++; w < n; w = next++) body(w);
    });
  };
  {
    TraceRange tr("sba sort");
    over_windows([&](int w) { rc[w] = sort_window(h, problems[w], w, inner, sorts[w], win[w], err[w]); });
  }
  for (int w = 0; w < n; ++w)
    if (rc[w] != SBA_OK) return fail(h, rc[w], err[w]);
  int64_t obs = 0, pairs = 0, points = 0, poses = 0, cams = 0, dups = 0;
  for (int w = 0; w < n; ++w) {
    WindowDesc& W = win[w];
    dups += sorts[w].ndup;
    W.obs_off = static_cast<int32_t>(obs);
    W.pair_off = static_cast<int32_t>(pairs);
    W.point_off = static_cast<int32_t>(points);
    W.pose_off = static_cast<int32_t>(poses);
    W.cam_off = static_cast<int32_t>(cams);
    obs += W.no;
    pairs += sorts[w].npairs;
    points += W.np;
    poses += W.nc;
    cams += W.ncam;
  }
  if (obs > INT32_MAX || pairs + dups > INT32_MAX)
    return fail(h, SBA_E_CAPACITY, "more than 2^31 observations in a batch");
  // the wide points' pair copies follow the pairs of all windows
  for (int w = 0, d = 0; w < n; ++w) {
    sorts[w].dup_off = pairs + d;
    d += static_cast<int>(sorts[w].ndup);
  }
  H.obs_orig.resize(obs);
  H.obs_xy.resize(obs);
  H.obs_info.resize(obs);
  H.pair_obs.resize(pairs + dups + 1);
  H.pair_pose.resize(pairs + dups);
  H.pair_rec.resize(pairs + dups);
  H.point_pair.resize(points + 1);
  H.point_nfree.resize(points);
  H.pts.resize(points);
  H.poses.resize(3 * poses);
  H.cams.resize(3 * cams);
  H.pair_obs[pairs + dups] = static_cast<int32_t>(obs);  // end sentinels
  H.point_pair[points] = static_cast<int32_t>(pairs);
  {
    TraceRange tr("sba fill");
    over_windows([&](int w) { fill_window(h, problems[w], w, inner, sorts[w], win[w], H, tiles[w], btiles[w]); });
  }
  for (int w = 0; w < n; ++w) {
    H.tiles.insert(H.tiles.end(), tiles[w].begin(), tiles[w].end());
    H.btiles.insert(H.btiles.end(), btiles[w].begin(), btiles[w].end());
  }
  return SBA_OK;
}

template <class T>
int upload(sba_handle* h, sba_batch* b, const std::vector<T>& v, T** out, size_t min_count = 1) {
  const size_t count = std::max(v.size(), min_count);
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, count * sizeof(T)));
  b->allocs.push_back(p);
  if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<T*>(p);
  return SBA_OK;
}

template <class T>
int alloc(sba_handle* h, sba_batch* b, size_t count, T** out) {
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  b->allocs.push_back(p);
  HIP_TRY(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
  *out = static_cast<T*>(p);
  return SBA_OK;
}

int free_batch(sba_batch* b) {
  if (!b) return SBA_OK;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : b->allocs) (void)hipFree(p);
  delete b;
  return SBA_OK;
}

// Device memory of a batch: every stream at a 256-byte aligned offset of one block. The uploaded streams come
// first, then the streams that start at zero, which one fill kernel clears (no memset node, so the sequence can
// sit inside a caller's stream capture).
struct Layout {
  struct Item {
    size_t off, bytes;
    const void* host;
    size_t host_bytes;
    std::function<void(char*)> bind;
  };
  std::vector<Item> items;
  size_t total = 0, zero_begin = 0;

  void add(size_t bytes, const void* host, size_t host_bytes, std::function<void(char*)> bind) {
    items.push_back(Item{total, bytes, host, host_bytes, std::move(bind)});
    total = (total + std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
  }
  template <class P, class T, class A>
  void upload(P& field, const std::vector<T, A>& v) {
    add(std::max<size_t>(v.size(), 1) * sizeof(T), v.empty() ? nullptr : v.data(), v.size() * sizeof(T),
        [&field](char* p) { field = reinterpret_cast<P>(p); });
  }
  void begin_zeroed() { zero_begin = total; }
  template <class P>
  void zeroed(P& field, size_t count) {
    add(std::max<size_t>(count, 1) * sizeof(*field), nullptr, 0, [&field](char* p) { field = reinterpret_cast<P>(p); });
  }
};

// Where a batch's device block comes from: its own allocation, the handle's workspace, or the scratch array.
using MemorySource = std::function<int(sba_batch* b, size_t bytes, char** base)>;

// Binds the layout to `base`, uploads the host streams and clears the zeroed ones on the handle's stream; with
// `wait`, waits for the stream. Without it the copies may still read the handle's packing buffers when this
// returns: the next packing waits for the stream before it writes them.
int place(sba_handle* h, const Layout& L, char* base, bool wait) {
  TraceRange tr("sba upload");
  for (const auto& it : L.items) it.bind(base + it.off);
  for (const auto& it : L.items)
    if (it.host_bytes) HIP_TRY(hipMemcpyAsync(base + it.off, it.host, it.host_bytes, hipMemcpyHostToDevice, h->stream));
  if (L.total > L.zero_begin) HIP_TRY(launch_fill_zero(base + L.zero_begin, L.total - L.zero_begin, h->stream));
  if (wait) HIP_TRY(hipStreamSynchronize(h->stream));
  return SBA_OK;
}

// Every stream the kernels dereference must be bound before the first launch (a null stream would fault the GPU).
int check_streams(sba_handle* h, const Batch& d) {
  const void* need[] = {d.win, d.st, d.tiles, d.chunks, d.btiles, d.bchunks, d.red_first, d.obs_xy, d.obs_info,
                        d.pair_obs, d.pair_pose, d.pair_rec, d.pair_xy, d.pair_info, d.point_pair, d.point_nfree,
                        d.pts0, d.pts1, d.pts_init, d.paux, d.poses0, d.poses1, d.poses_init, d.cams, d.chunk_map,
                        d.gfirst, d.gcon, d.part, d.sred, d.A, d.dc, d.dc64, d.out_pts, d.out_poses, d.trace, d.bpart};
  for (size_t k = 0; k < sizeof(need) / sizeof(need[0]); ++k)
    if (!need[k]) return fail(h, SBA_E_DEVICE, "internal error: device stream " + std::to_string(k) + " is not bound");
  return SBA_OK;
}

// Pairs of work per chunk when the options leave it open: about 32 chunks per compute unit, so the hardware
// dispatcher balances uneven windows, but at least 1024 pairs so that a chunk's record stays small against its
// work (measured on the 512-window batch: 3000 pairs per chunk is within 2% of the best of 1000..25000).
int64_t default_chunk_pairs(int64_t total_pairs) {
  return std::min<int64_t>(std::max<int64_t>(1024, total_pairs / 8192), 65536);
}

// The launch sequence of one solve. With events, each kernel class is timed (eager mode only).
int enqueue_solve(sba_handle* h, sba_batch* b, bool timed) {
  const Batch& d = b->dev;
  hipStream_t s = h->stream;
  HIP_TRY(launch_reset(d, s));
  HIP_TRY(launch_backsub(d, true, s));
  HIP_TRY(launch_decide(d, true, s));
  for (int it = 0; it < b->max_iter; ++it) {
    if (timed) HIP_TRY(hipEventRecord(b->ev[0], s));
    HIP_TRY(launch_schur(d, s));
    if (timed) HIP_TRY(hipEventRecord(b->ev[1], s));
    HIP_TRY(launch_reduce(d, s));
    if (timed) HIP_TRY(hipEventRecord(b->ev[2], s));
    HIP_TRY(launch_factor(d, nullptr, s));
    for (int w : b->mapped) HIP_TRY(launch_large_factor(d, w, 6 * b->win[w].ncf, s));
    if (timed) HIP_TRY(hipEventRecord(b->ev[3], s));
    HIP_TRY(launch_backsub(d, false, s));
    if (timed) HIP_TRY(hipEventRecord(b->ev[4], s));
    HIP_TRY(launch_decide(d, false, s));
    if (timed) {
      HIP_TRY(hipEventRecord(b->ev[5], s));
      HIP_TRY(hipEventSynchronize(b->ev[5]));
      for (int k = 0; k < 5; ++k) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev[k], b->ev[k + 1]));
        b->class_ms[k] += ms;
        b->class_launches[k] += 1;
      }
    }
  }
  HIP_TRY(launch_gather(d, s));
  return SBA_OK;
}

}  // namespace

extern "C" {

int sba_abi_version(void) { return SBA_ABI_VERSION; }

void sba_default_options(sba_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->device = 0;
  o->loss = SBA_LOSS_STUDENT_T;
  o->fixed_iterations = 0;
  o->use_graph = 1;
  o->chunk_pairs = 0;
}

int sba_create(const sba_options* opt, sba_handle** out) {
  if (!out) return SBA_E_INVALID;
  *out = nullptr;
  sba_handle* h = new (std::nothrow) sba_handle;
  if (!h) return SBA_E_INVALID;
  if (opt)
    h->opt = *opt;
  else
    sba_default_options(&h->opt);
  if (h->opt.loss != SBA_LOSS_STUDENT_T && h->opt.loss != SBA_LOSS_HUBER) {
    delete h;
    return SBA_E_INVALID;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= h->opt.device || h->opt.device < 0) {
    delete h;
    return SBA_E_DEVICE;
  }
  if (hipSetDevice(h->opt.device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return SBA_E_DEVICE;
  }
  (void)set_kernel_attributes();  // raises the dynamic shared-memory limit where the runtime needs it
  void* scratch = nullptr;
  if (scratch_setup(h->opt.device, &scratch, &h->scratch_bytes) != 0) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return SBA_E_DEVICE;
  }
  h->scratch = static_cast<char*>(scratch);
  *out = h;
  return SBA_OK;
}

int sba_destroy(sba_handle* h) {
  if (!h) return SBA_OK;
  (void)hipSetDevice(h->opt.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->ws) (void)hipFree(h->ws);
  if (h->imu_ws) (void)hipFree(h->imu_ws);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SBA_OK;
}

const char* sba_last_error(const sba_handle* h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"

namespace sba_internal {

hipStream_t stream_of(sba_handle* h) { return h->stream; }
void parallel_for(sba_handle* h, int n, int threads, const std::function<void(int)>& fn) {
  const int t = std::max(1, std::min(threads, n));
  h->pool.run(t, [&](int k) {
    for (int i = k; i < n; i += t) fn(i);
  });
}
int device_of(const sba_handle* h) { return h->opt.device; }
int set_error(sba_handle* h, int code, const std::string& msg) { return fail(h, code, msg); }

int aux_workspace(sba_handle* h, size_t bytes, char** out) {
  if (bytes > h->imu_ws_bytes) {
    if (h->imu_ws) {
      HIP_TRY(hipStreamSynchronize(h->stream));
      HIP_TRY(hipFree(h->imu_ws));
      h->imu_ws = nullptr;
      h->imu_ws_bytes = 0;
    }
    const size_t grown = std::max(bytes, 2 * h->imu_ws_bytes);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&h->imu_ws), grown));
    h->imu_ws_bytes = grown;
  }
  *out = h->imu_ws;
  return SBA_OK;
}

}  // namespace sba_internal

namespace {

// Packs and plans the windows on the host, takes the batch's device block from `memory`, uploads, and (graph)
// records the whole solve as one hipGraph. On failure nothing is left allocated by the batch itself.
int build_batch_impl(sba_handle* h, const sba_problem_view* problems, int32_t n, bool graph, bool wait_upload,
                     const MemorySource& memory, sba_batch** out) {
  *out = nullptr;
  TraceRange tr("sba pack and plan");
  HostPack& H = h->pack;
  HIP_TRY(hipStreamSynchronize(h->stream));  // no copy of an earlier batch still reads the packing buffers
  H.clear();
  std::unique_ptr<sba_batch, int (*)(sba_batch*)> owned(new (std::nothrow) sba_batch, free_batch);
  sba_batch* b = owned.get();
  if (!b) return fail(h, SBA_E_INVALID, "out of host memory");
  b->n = n;
  b->win.resize(n);
  b->status_pre.assign(n, SBA_OK);
  int64_t sred_off = 0, a_off = 0, dc_off = 0;
  int max_nc = 1, max_ncam = 1, max_ncf = 1, max_rec = 0;
  int rc = SBA_OK;
  if ((rc = pack_windows(h, problems, n, H, b->win)) != SBA_OK) return rc;
  for (int w = 0; w < n; ++w) {
    const sba_problem_view& p = problems[w];
    WindowDesc& W = b->win[w];
    if (W.np == 0 || W.nc == 0 || W.no == 0) b->status_pre[w] = SBA_E_EMPTY;
    W.nt = W.mapped ? 0 : (6 * W.ncf + 15) / 16;
    W.rec = W.mapped ? 0 : 256 * (W.nt * (W.nt + 1) / 2) + 33 * W.ncf;
    W.sred_off = sred_off;
    sred_off += W.rec;
    max_rec = std::max(max_rec, W.rec);
    W.dc_off = static_cast<int32_t>(dc_off);
    dc_off += 6 * W.ncf;
    const int nn = 6 * W.ncf;
    W.a_off = 0;
    if (W.mapped) {
      // FP64 A (n x n), rhs, diag U, the inverses of the diagonal blocks
      W.a_off = a_off;
      a_off += static_cast<int64_t>(nn) * nn + 2 * nn + static_cast<int64_t>((nn + 63) / 64) * 64 * 64;
      b->mapped.push_back(w);
    }
    if (b->point_order.size() != static_cast<size_t>(n)) b->point_order.resize(n);
    b->point_order[w] = H.sorts[w].order;
    max_nc = std::max(max_nc, W.nc);
    max_ncam = std::max(max_ncam, W.ncam);
    max_ncf = std::max(max_ncf, W.ncf);
    b->max_iter = std::max(b->max_iter, p.max_iterations);
    // an empty window never runs: the initial decision sees zero observations as infeasible, so mark it here
    if (b->status_pre[w] != SBA_OK) W.max_iterations = 0;
  }
  b->total_obs = static_cast<int64_t>(H.obs_xy.size());
  b->total_pairs = static_cast<int64_t>(H.pair_pose.size());

  // chunks: consecutive tiles of one window, about chunk_pairs pairs of work each; each owns one record
  int64_t target = h->opt.chunk_pairs;
  if (target <= 0) target = default_chunk_pairs(b->total_pairs);
  const int64_t btarget = h->opt.bchunk_pairs > 0 ? h->opt.bchunk_pairs : target;
  // In a mapped window a chunk also stages at most kMaxFreePoses free key frames: its map lists them in
  // increasing order and its pairs carry their index in that list.
  pinned_vector<Chunk>& chunks = H.chunks;
  pinned_vector<int32_t>& chunk_map = H.chunk_map;
  int64_t part_off = 0;
  {
    size_t t = 0;
    int last_w = -1;
    std::vector<int> slot;
    while (t < H.tiles.size()) {
      const int w = H.tiles[t].window;
      WindowDesc& W = b->win[w];
      if (w != last_w) {
        W.part_off = part_off;
        W.nchunks = 0;
        W.chunk_begin = static_cast<int32_t>(chunks.size());
        last_w = w;
        if (W.mapped) slot.assign(W.ncf, -1);
      }
      const size_t begin = t;
      int64_t pairs = 0;
      std::vector<int> kfs;  // free key frames of the chunk (mapped windows)
      // a wide point's tile is a chunk of its own
      while (t < H.tiles.size() && H.tiles[t].window == w &&
             (t == begin || (pairs < target && !H.tiles[t].virt && !H.tiles[begin].virt))) {
        const Tile& T = H.tiles[t];
        if (W.mapped) {
          std::vector<int> add;
          for (int q = T.pair_begin; q < T.pair_end; ++q) {
            const int j = H.pair_pose[q];
            if (j < W.ncf && slot[j] < 0 && std::find(add.begin(), add.end(), j) == add.end()) add.push_back(j);
          }
          if (t > begin && kfs.size() + add.size() > static_cast<size_t>(kMaxFreePoses)) break;
          for (int j : add) {
            slot[j] = 0;
            kfs.push_back(j);
          }
        }
        pairs += T.pair_end - T.pair_begin;
        ++t;
      }
      Chunk c{static_cast<int32_t>(begin), static_cast<int32_t>(t), part_off, W.ncf, W.nt, 0, w};
      int64_t rec = W.rec;
      if (W.mapped) {
        std::sort(kfs.begin(), kfs.end());
        for (size_t k = 0; k < kfs.size(); ++k) slot[kfs[k]] = static_cast<int>(k);
        for (size_t tt = begin; tt < t; ++tt)
          for (int q = H.tiles[tt].pair_begin; q < H.tiles[tt].pair_end; ++q) {
            const int j = H.pair_pose[q];
            if (j < W.ncf) H.pair_rec[q].w = (H.pair_rec[q].w & 255) | (slot[j] << 8);
          }
        for (int j : kfs) slot[j] = -1;
        c.nloc = static_cast<int32_t>(kfs.size());
        c.nt = (6 * c.nloc + 15) / 16;
        c.map_off = static_cast<int32_t>(chunk_map.size());
        chunk_map.insert(chunk_map.end(), kfs.begin(), kfs.end());
        rec = 36 * (c.nloc * (c.nloc + 1) / 2) + 33 * c.nloc;  // key-frame blocks, then U, e, diag U
      }
      chunks.push_back(c);
      part_off += rec;
      W.nchunks += 1;
    }
  }
  // Mapped windows: the contributions to each key-frame block (a, b), a <= b, of the reduced system, listed
  // chunk by chunk (counting sort over the blocks), so that the gather sums every block in chunk order. A chunk
  // of a wide point contributes the blocks between its two groups, and those inside a group it keeps.
  auto contributes = [&](const Chunk& c, int la, int lb) {
    const int v = H.tiles[c.tile_begin].virt;
    if (!v) return true;
    const int na = v >> 8;
    if (lb < na) return (v & 2) != 0;
    if (la >= na) return (v & 4) != 0;
    return true;
  };
  {
    pinned_vector<int32_t>& gfirst = H.gfirst;
    pinned_vector<int4>& gcon = H.gcon;
    std::vector<int32_t> cur;
    for (int w : b->mapped) {
      WindowDesc& W = b->win[w];
      const int ncf = W.ncf;
      const size_t nb = static_cast<size_t>(ncf) * ncf, base = gfirst.size();
      W.gfirst_off = static_cast<int32_t>(base);
      gfirst.resize(base + nb + 1, 0);
      int32_t* f = gfirst.data() + base;
      const int c0 = W.chunk_begin, c1 = W.chunk_begin + W.nchunks;
      for (int ci = c0; ci < c1; ++ci) {
        const int32_t* m = chunk_map.data() + chunks[ci].map_off;
        for (int la = 0; la < chunks[ci].nloc; ++la)
          for (int lb = la; lb < chunks[ci].nloc; ++lb)
            if (contributes(chunks[ci], la, lb)) ++f[static_cast<size_t>(m[la]) * ncf + m[lb] + 1];
      }
      f[0] = static_cast<int32_t>(gcon.size());
      for (size_t i = 1; i <= nb; ++i) f[i] += f[i - 1];
      cur.assign(f, f + nb);
      gcon.resize(static_cast<size_t>(f[nb]));
      for (int ci = c0; ci < c1; ++ci) {
        const int32_t* m = chunk_map.data() + chunks[ci].map_off;
        for (int la = 0; la < chunks[ci].nloc; ++la)
          for (int lb = la; lb < chunks[ci].nloc; ++lb)
            if (contributes(chunks[ci], la, lb))
              gcon[cur[static_cast<size_t>(m[la]) * ncf + m[lb]]++] =
                  make_int4(ci, la | (lb << 8) | (chunks[ci].nt << 16) | (chunks[ci].nloc << 24),
                            static_cast<int32_t>(chunks[ci].part_off & 0xffffffff),
                            static_cast<int32_t>(chunks[ci].part_off >> 32));
      }
    }
  }
  // back-substitution chunks: the same grouping over the larger back-substitution tiles
  pinned_vector<Chunk>& bchunks = H.bchunks;
  {
    size_t t = 0;
    while (t < H.btiles.size()) {
      const int w = H.btiles[t].window;
      const size_t begin = t;
      int64_t pairs = 0;
      while (t < H.btiles.size() && H.btiles[t].window == w && (t == begin || pairs < btarget)) {
        pairs += H.btiles[t].pair_end - H.btiles[t].pair_begin;
        ++t;
      }
      WindowDesc& W = b->win[w];
      if (W.nbchunks == 0) W.bchunk_begin = static_cast<int32_t>(bchunks.size());
      W.nbchunks += 1;
      bchunks.push_back(Chunk{static_cast<int32_t>(begin), static_cast<int32_t>(t), 0, 0, 0, 0, w});
    }
  }

  // record reduction work list: 64 entries per workgroup for windows of many chunks, 256 otherwise
  pinned_vector<int32_t>& red_first = H.red_first;
  red_first.assign(n + 1, 0);
  for (int w = 0; w < n; ++w) {
    const WindowDesc& W = b->win[w];
    const int per = W.nchunks < 64 ? 256 : 64;
    red_first[w + 1] = red_first[w] + (W.mapped ? 0 : (W.rec + per - 1) / per);
  }

  Batch& d = b->dev;
  d.n_windows = n;
  d.red_blocks = red_first[n];
  d.n_chunks = static_cast<int32_t>(chunks.size());
  d.n_bchunks = static_cast<int32_t>(bchunks.size());
  d.n_points = static_cast<int32_t>(H.pts.size());
  d.n_poses = static_cast<int32_t>(H.poses.size() / 3);
  d.n_cams = static_cast<int32_t>(H.cams.size() / 3);
  d.loss = h->opt.loss;
  d.fixed_iterations = h->opt.fixed_iterations;
  d.max_rec = max_rec;
  d.max_nc = max_nc;
  Layout L;
  L.upload(d.win, b->win);
  L.upload(d.tiles, H.tiles);
  L.upload(d.chunks, chunks);
  L.upload(d.chunk_map, chunk_map);
  L.upload(d.gfirst, H.gfirst);
  L.upload(d.gcon, H.gcon);
  L.upload(d.btiles, H.btiles);
  L.upload(d.bchunks, bchunks);
  L.upload(d.red_first, red_first);
  L.upload(d.obs_xy, H.obs_xy);
  L.upload(d.obs_info, H.obs_info);
  L.upload(d.pair_obs, H.pair_obs);
  L.upload(d.pair_pose, H.pair_pose);
  L.upload(d.pair_rec, H.pair_rec);
  bool copies = false;  // wide points' pair copies
  for (int w = 0; w < n; ++w) copies = copies || H.sorts[w].ndup > 0;
  const bool pair_streams = H.obs_xy.size() != H.pair_rec.size() || copies;
  if (pair_streams) {
    // some pair has several observations: the first observation of every pair, indexed by pair
    H.pair_xy.resize(H.pair_rec.size());
    H.pair_info.resize(H.pair_rec.size());
    for (size_t q = 0; q < H.pair_rec.size(); ++q) {
      H.pair_xy[q] = H.obs_xy[H.pair_rec[q].x];
      H.pair_info[q] = H.obs_info[H.pair_rec[q].x];
    }
    L.upload(d.pair_xy, H.pair_xy);
    L.upload(d.pair_info, H.pair_info);
  }
  L.upload(d.point_pair, H.point_pair);
  L.upload(d.point_nfree, H.point_nfree);
  L.upload(d.pts_init, H.pts);
  L.upload(d.poses_init, H.poses);
  L.upload(d.cams, H.cams);
  L.upload(b->obs_orig, H.obs_orig);
  L.begin_zeroed();
  L.zeroed(d.st, static_cast<size_t>(n));
  L.zeroed(d.pts0, H.pts.size());
  L.zeroed(d.pts1, H.pts.size());
  L.zeroed(d.paux, 3 * H.pts.size());
  L.zeroed(d.poses0, H.poses.size());
  L.zeroed(d.poses1, H.poses.size());
  L.zeroed(d.part, static_cast<size_t>(part_off));
  L.zeroed(d.sred, static_cast<size_t>(sred_off));
  L.zeroed(d.A, static_cast<size_t>(a_off));
  L.zeroed(d.dc, static_cast<size_t>(dc_off));
  L.zeroed(d.dc64, static_cast<size_t>(dc_off));
  L.zeroed(d.out_pts, H.pts.size());
  L.zeroed(d.out_poses, H.poses.size());
  L.zeroed(d.trace, static_cast<size_t>(n) * kTrace);
  L.zeroed(d.bpart, bchunks.size());
  d.dbg_dp = nullptr;
  // shared-memory layout of the Schur kernel, sized for the batch's largest window. The key frames are staged
  // only while the layout leaves room for two workgroups per compute unit (80 KiB each); otherwise (large
  // windows) the kernel reads them from global memory (lds_pose = -1).
  int off = 0;
  const int fixed_words = 12 * max_ncam + 12 * kTilePoints + kTilePoints * kMaxFreePoses + 24 * (kTilePairs + 1) +
                          2 * 4 * kTilePoints + kKt * kZcLd + 128;
  const bool stage_poses = (fixed_words + 12 * max_nc) * 4 + 64 <= 80 * 1024;
  d.lds_pose = stage_poses ? off : -1;
  if (stage_poses) off += 12 * max_nc;
  d.lds_cam = off;
  off += 12 * max_ncam;
  d.lds_acc = off;
  off += 12 * kTilePoints;
  d.lds_ent = off;
  off += kTilePoints * kMaxFreePoses;
  d.lds_ec = off;
  off += 6 * (kTilePairs + 1);
  d.lds_w = off;
  off += 18 * (kTilePairs + 1);
  off = (off + 3) & ~3;  // float4 staging
  d.lds_pts = off;
  off += 2 * 4 * kTilePoints;
  off = (off + 63) & ~63;
  d.lds_zc = off;
  off += kKt * kZcLd;
  d.lds_words = off;
  if (static_cast<int64_t>(d.lds_words) * 4 + 64 > 160 * 1024)
    return fail(h, SBA_E_CAPACITY, "Schur partial does not fit in shared memory");
  char* base = nullptr;
  if ((rc = memory(b, L.total, &base)) != SBA_OK) return rc;
  if ((rc = place(h, L, base, wait_upload)) != SBA_OK) return rc;
  if (!pair_streams) {
    // every pair has exactly one observation: the observation streams are already in pair order
    d.pair_xy = d.obs_xy;
    d.pair_info = d.obs_info;
  }
  if ((rc = check_streams(h, d)) != SBA_OK) return rc;
  if (graph) {
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int erc = enqueue_solve(h, b, false);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (erc != SBA_OK || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(h, SBA_E_DEVICE, "graph capture failed");
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  }
  *out = owned.release();
  return SBA_OK;
}

// Packs, uploads and (with `graph`) captures a batch; host allocation failures (the packing buffers are
// page-locked) come back as a status, not as an exception through the C ABI.
int build_batch(sba_handle* h, const sba_problem_view* problems, int32_t n, bool graph, const MemorySource& memory,
                sba_batch** out, bool wait_upload = true) {
  try {
    return build_batch_impl(h, problems, n, graph, wait_upload, memory, out);
  } catch (const std::bad_alloc&) {
    *out = nullptr;
    return fail(h, SBA_E_INVALID, "out of host memory while packing");
  }
}

// One window on the handle's stream, its device block taken from `memory`: packed, uploaded, solved with eager
// launches (a different window shape on every call would need a new graph each time) and read back in place.
int solve_one(sba_handle* h, const sba_problem_view* problem, sba_result* result, const MemorySource& memory) {
  if (!h || !problem || !result) return fail(h, SBA_E_INVALID, "bad arguments to sba_solve");
  if (problem->num_points == 0 || problem->num_poses == 0 || problem->num_observations == 0) {
    result->status = SBA_E_EMPTY;
    result->iterations = 0;
    return SBA_E_EMPTY;
  }
  HIP_TRY(hipSetDevice(h->opt.device));
  TraceRange tr("SchurComplementBundlerGpu::solve");
  sba_batch* b = nullptr;
  int rc = build_batch(h, problem, 1, false, memory, &b, false);  // sba_batch_read below waits for the stream
  if (rc != SBA_OK) return rc;
  std::unique_ptr<sba_batch, int (*)(sba_batch*)> owned(b, free_batch);
  if ((rc = enqueue_solve(h, b, false)) != SBA_OK) return rc;
  if ((rc = sba_batch_read(h, b, result)) != SBA_OK) return rc;
  return result->status;
}

// A device block of its own for each batch (sba_batch_create).
int own_memory(sba_handle* h, sba_batch* b, size_t bytes, char** base) {
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, bytes));
  b->allocs.push_back(p);
  *base = static_cast<char*>(p);
  return SBA_OK;
}

// The handle's workspace, regrown to twice the need when a batch does not fit (sba_solve).
int workspace_memory(sba_handle* h, size_t bytes, char** base) {
  if (bytes > h->ws_bytes) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->ws) HIP_TRY(hipFree(h->ws));
    h->ws = nullptr;
    h->ws_bytes = 0;
    const size_t grown = std::max(bytes, 2 * h->ws_bytes);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, grown));
    h->ws = static_cast<char*>(p);
    h->ws_bytes = grown;
  }
  *base = h->ws;
  return SBA_OK;
}

}  // namespace

extern "C" {

int sba_batch_create(sba_handle* h, const sba_problem_view* problems, int32_t n, sba_batch** out) {
  if (!h || !problems || n <= 0 || !out) return fail(h, SBA_E_INVALID, "bad arguments to sba_batch_create");
  HIP_TRY(hipSetDevice(h->opt.device));
  return build_batch(h, problems, n, h->opt.use_graph != 0,
                     [h](sba_batch* b, size_t bytes, char** base) { return own_memory(h, b, bytes, base); }, out);
}

int sba_batch_run(sba_handle* h, sba_batch* b) {
  if (!h || !b) return SBA_E_INVALID;
  TraceRange tr("sba_batch_run");
  HIP_TRY(hipSetDevice(h->opt.device));
  b->adj_ready = false;
  if (b->gexec) {
    HIP_TRY(hipGraphLaunch(b->gexec, h->stream));
    return SBA_OK;
  }
  return enqueue_solve(h, b, false);
}

int sba_batch_sync(sba_handle* h) {
  if (!h) return SBA_E_INVALID;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return SBA_OK;
}

int sba_batch_read(sba_handle* h, sba_batch* b, sba_result* results) {
  if (!h || !b || !results) return SBA_E_INVALID;
  TraceRange tr("sba_batch_read");
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::vector<WindowState> st(b->n);
  HIP_TRY(hipMemcpy(st.data(), b->dev.st, sizeof(WindowState) * b->n, hipMemcpyDeviceToHost));
  std::vector<float4> pts(b->dev.n_points), poses(3 * static_cast<size_t>(b->dev.n_poses));
  HIP_TRY(hipMemcpy(pts.data(), b->dev.out_pts, sizeof(float4) * pts.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(poses.data(), b->dev.out_poses, sizeof(float4) * poses.size(), hipMemcpyDeviceToHost));
  for (int w = 0; w < b->n; ++w) {
    const WindowDesc& W = b->win[w];
    sba_result& r = results[w];
    const WindowState& s = st[w];
    r.iterations = s.iterations;
    r.initial_cost = static_cast<float>(s.initial_cost);
    r.last_cost = static_cast<float>(s.last_cost);
    r.initial_cost_f64 = s.initial_cost;
    r.last_cost_f64 = s.last_cost;
    r.lambda = s.lambda;
    r.accepted = s.accepted;
    r.status = b->status_pre[w] != SBA_OK ? b->status_pre[w] : s.status;
    if (r.status == SBA_E_EMPTY) {
      r.iterations = 0;
      r.initial_cost = r.last_cost = 0.f;
      r.initial_cost_f64 = r.last_cost_f64 = 0.0;
    }
    // state is written back only for a solve that ran (the reference leaves the problem untouched on failure)
    if (r.status != SBA_OK) continue;
    if (r.points) {
      const std::vector<int32_t>& order = b->point_order[w];
      for (int i = 0; i < W.np; ++i) {
        const float4 q = pts[W.point_off + i];
        const int o = order.empty() ? i : order[i];
        r.points[3 * o] = q.x;
        r.points[3 * o + 1] = q.y;
        r.points[3 * o + 2] = q.z;
      }
    }
    if (r.rig_from_world)
      for (int j = 0; j < W.nc; ++j) {
        float* m = r.rig_from_world + 16 * j;
        const float4* rows = &poses[3 * (W.pose_off + j)];
        const float R[3][4] = {{rows[0].x, rows[0].y, rows[0].z, rows[0].w},
                               {rows[1].x, rows[1].y, rows[1].z, rows[1].w},
                               {rows[2].x, rows[2].y, rows[2].z, rows[2].w}};
        for (int c = 0; c < 4; ++c)
          for (int rr = 0; rr < 3; ++rr) m[4 * c + rr] = R[rr][c];
        m[3] = m[7] = m[11] = 0.f;
        m[15] = 1.f;
      }
  }
  return SBA_OK;
}

int sba_batch_trace(sba_handle* h, sba_batch* b, int32_t window, int32_t capacity, double* rho, double* trial_cost,
                    float* lambda, int32_t* flags, int32_t* n) {
  if (!h || !b || window < 0 || window >= b->n || capacity < 0) return fail(h, SBA_E_INVALID, "bad sba_batch_trace arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  WindowState st;
  std::vector<TraceEntry> t(kTrace);
  HIP_TRY(hipMemcpy(&st, b->dev.st + window, sizeof(st), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(t.data(), b->dev.trace + static_cast<int64_t>(window) * kTrace, sizeof(TraceEntry) * kTrace,
                    hipMemcpyDeviceToHost));
  const int m = std::min({st.iterations, kTrace, capacity});
  for (int k = 0; k < m; ++k) {
    if (rho) rho[k] = t[k].rho;
    if (trial_cost) trial_cost[k] = t[k].trial_cost;
    if (lambda) lambda[k] = t[k].lambda;
    if (flags) flags[k] = t[k].flags;
  }
  if (n) *n = m;
  return SBA_OK;
}

int sba_batch_destroy(sba_handle* h, sba_batch* b) {
  if (h) (void)hipSetDevice(h->opt.device);
  return free_batch(b);
}

int sba_batch_info(sba_handle* h, sba_batch* b, int64_t* num_observations, int64_t* num_pairs, int32_t* num_chunks,
                   int32_t* num_windows) {
  if (!h || !b) return SBA_E_INVALID;
  if (num_observations) *num_observations = b->total_obs;
  if (num_pairs) *num_pairs = b->total_pairs;
  if (num_chunks) *num_chunks = b->dev.n_chunks;
  if (num_windows) *num_windows = b->n;
  return SBA_OK;
}

int sba_batch_kernel_times(sba_handle* h, sba_batch* b, float* ms_per_class, int32_t* launches_per_class) {
  if (!h || !b) return SBA_E_INVALID;
  for (int k = 0; k < 5; ++k) {
    b->class_ms[k] = 0.f;
    b->class_launches[k] = 0;
  }
  for (auto& e : b->ev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  const int rc = enqueue_solve(h, b, true);
  if (rc != SBA_OK) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int k = 0; k < 5; ++k) {
    if (ms_per_class) ms_per_class[k] = b->class_ms[k];
    if (launches_per_class) launches_per_class[k] = b->class_launches[k];
  }
  return SBA_OK;
}

int sba_batch_winsorize(sba_handle* h, sba_batch* b, int32_t window, int32_t key_frame, uint8_t* reset_mask,
                        int32_t* n_reset) {
  if (!h || !b || window < 0 || window >= b->n) return fail(h, SBA_E_INVALID, "bad sba_batch_winsorize arguments");
  const WindowDesc& W = b->win[window];
  if (key_frame < 0 || key_frame >= W.nc) return fail(h, SBA_E_INVALID, "key frame out of range");
  HIP_TRY(hipSetDevice(h->opt.device));
  hipStream_t s = h->stream;
  const size_t no = static_cast<size_t>(std::max(W.no, 1)), np = static_cast<size_t>(std::max(W.np, 1));
  // workspace: squared errors (FP64) and point index per observation, counters, marks per point
  const size_t off_pt = 8 * no, off_cnt = (off_pt + 4 * no + 255) & ~size_t(255), off_mark = off_cnt + 256;
  char* ws = nullptr;
  HIP_TRY(hipMalloc(&ws, off_mark + np));
  auto done = [&](int rc) {
    (void)hipFree(ws);
    return rc;
  };
  int32_t out[4] = {0, 0, 0, 0};
  int32_t* cnt = reinterpret_cast<int32_t*>(ws + off_cnt);
  uint8_t* mark = reinterpret_cast<uint8_t*>(ws + off_mark);
  if (hipMemsetAsync(ws + off_cnt, 0, 256, s) != hipSuccess || hipMemsetAsync(mark, 0, np, s) != hipSuccess ||
      launch_winsorize_select(b->dev, window, W.np, key_frame, reinterpret_cast<double*>(ws),
                              reinterpret_cast<int32_t*>(ws + off_pt), cnt, mark, cnt + 4, s) != hipSuccess ||
      hipMemcpyAsync(out, cnt + 4, sizeof(out), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(h, SBA_E_DEVICE, "winsorize launch failed"));
  if (out[3]) return done(fail(h, SBA_E_CAPACITY, "more than 1024 observations share the winsorizing cut"));
  // the marks are by packed point: a window whose points were reordered maps them back
  const std::vector<int32_t>& order = b->point_order[window];
  std::vector<uint8_t> packed_mark(reset_mask && !order.empty() ? W.np : 0);
  uint8_t* mark_dst = order.empty() ? reset_mask : packed_mark.data();
  if (launch_winsorize_apply(b->dev, window, W.np, W.nc, mark, s) != hipSuccess ||
      (reset_mask && W.np > 0 && hipMemcpyAsync(mark_dst, mark, W.np, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(h, SBA_E_DEVICE, "winsorize apply failed"));
  if (reset_mask && !order.empty())
    for (int i = 0; i < W.np; ++i) reset_mask[order[i]] = packed_mark[i];
  b->adj_ready = false;
  if (n_reset) *n_reset = out[0];
  return done(SBA_OK);
}

int sba_batch_debug_step(sba_handle* h, sba_batch* b, double lambda, double* S, double* e, double* dc, double* dp,
                         double* scalars) {
  if (!h || !b) return SBA_E_INVALID;
  if (b->win[0].mapped) return fail(h, SBA_E_CAPACITY, "the phase-level step covers windows of at most 40 free key frames");
  HIP_TRY(hipSetDevice(h->opt.device));
  b->adj_ready = false;
  Batch d = b->dev;
  const WindowDesc& W = b->win[0];
  hipStream_t s = h->stream;
  if (!b->dbg_S) {
    HIP_TRY(hipMalloc(&b->dbg_S, sizeof(double) * (240 * 240 + 240)));
    b->allocs.push_back(b->dbg_S);
  }
  float4* dpbuf = nullptr;
  HIP_TRY(hipMalloc(&dpbuf, sizeof(float4) * std::max(1, W.np)));
  d.dbg_dp = dpbuf;
  int rc = SBA_OK;
  do {
    if (launch_reset(d, s) != hipSuccess || launch_backsub(d, true, s) != hipSuccess || launch_decide(d, true, s) != hipSuccess) {
      rc = fail(h, SBA_E_DEVICE, "launch failed");
      break;
    }
    WindowState st0;
    if (hipMemcpyAsync(&st0, d.st, sizeof(st0), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(h, SBA_E_DEVICE, "state read failed");
      break;
    }
    const double current = st0.current_cost;
    st0.lambda = static_cast<float>(lambda);
    st0.done = 0;
    if (hipMemcpyAsync(d.st, &st0, sizeof(st0), hipMemcpyHostToDevice, s) != hipSuccess ||
        launch_schur(d, s) != hipSuccess || launch_reduce(d, s) != hipSuccess || launch_factor(d, b->dbg_S, s) != hipSuccess ||
        launch_backsub(d, false, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(h, SBA_E_DEVICE, "step failed");
      break;
    }
    WindowState st1;
    std::vector<CostPart> parts(std::max(1, W.nbchunks));
    if (W.nbchunks && hipMemcpy(parts.data(), d.bpart + W.bchunk_begin, sizeof(CostPart) * W.nbchunks,
                                hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(h, SBA_E_DEVICE, "readback failed");
      break;
    }
    CostPart tot{0.0, 0.0, 0.0, 0, 0.f};
    for (int k = 0; k < W.nbchunks; ++k) {
      tot.cost += parts[k].cost;
      tot.hess += parts[k].hess;
      tot.scal += parts[k].scal;
      tot.skipped += parts[k].skipped;
      tot.maxdp = std::max(tot.maxdp, parts[k].maxdp);
    }
    const int n = 6 * W.ncf;
    std::vector<double> Sbuf(static_cast<size_t>(n) * n + n), dcv(n);
    std::vector<float4> dpv(std::max(1, W.np));
    if (hipMemcpy(&st1, d.st, sizeof(st1), hipMemcpyDeviceToHost) != hipSuccess ||
        (n && hipMemcpy(Sbuf.data(), b->dbg_S, sizeof(double) * Sbuf.size(), hipMemcpyDeviceToHost) != hipSuccess) ||
        (n && hipMemcpy(dcv.data(), d.dc64 + W.dc_off, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) ||
        hipMemcpy(dpv.data(), dpbuf, sizeof(float4) * dpv.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(h, SBA_E_DEVICE, "readback failed");
      break;
    }
    if (S) std::copy(Sbuf.begin(), Sbuf.begin() + static_cast<size_t>(n) * n, S);
    if (e) std::copy(Sbuf.begin() + static_cast<size_t>(n) * n, Sbuf.end(), e);
    if (dc) std::copy(dcv.begin(), dcv.end(), dc);
    if (dp)
      for (int i = 0; i < W.npf; ++i) {
        dp[3 * i] = dpv[i].x;
        dp[3 * i + 1] = dpv[i].y;
        dp[3 * i + 2] = dpv[i].z;
      }
    if (scalars) {
      scalars[0] = tot.skipped == W.no ? INFINITY : tot.cost;
      scalars[1] = tot.skipped;
      scalars[2] = tot.hess;
      scalars[3] = st1.scal_pose + tot.scal;
      scalars[4] = tot.maxdp;
      scalars[5] = st1.maxdc;
      scalars[6] = current;
      scalars[7] = st1.chol_ok;
    }
    if (!st1.chol_ok) rc = SBA_E_INFEASIBLE;
  } while (false);
  (void)hipFree(dpbuf);
  return rc;
}

int sba_debug_factor(sba_handle* h, int32_t n, const double* A, const double* bvec, double* x) {
  if (!h || !A || !bvec || !x || n <= 0 || n % 6 != 0 || n > 6 * kMaxFreePoses) return fail(h, SBA_E_INVALID, "bad debug_factor size");
  HIP_TRY(hipSetDevice(h->opt.device));
  const int ncf = n / 6;
  const int nt = (n + 15) / 16, ntu = nt * (nt + 1) / 2;
  // a reduced record holding -A as the Schur tiles (so that U - sum Z Z^T = A), U = 0, e = b, diag U = 0
  std::vector<float> rec(256 * ntu + 33 * ncf, 0.f);
  for (int I = 0; I < nt; ++I)
    for (int J = I; J < nt; ++J) {
      const int tile = I * nt - (I * (I - 1)) / 2 + (J - I);
      for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 16; ++c) {
          const int i = 16 * I + r, j = 16 * J + c;
          if (i >= n || j >= n) continue;
          const double v = i >= j ? A[static_cast<int64_t>(i) * n + j] : A[static_cast<int64_t>(j) * n + i];
          rec[256 * tile + 64 * (r & 3) + 16 * (r >> 2) + c] = static_cast<float>(-v);
        }
    }
  for (int i = 0; i < n; ++i) rec[256 * ntu + 21 * ncf + i] = static_cast<float>(bvec[i]);
  sba_batch* b = new (std::nothrow) sba_batch;
  if (!b) return fail(h, SBA_E_INVALID, "out of host memory");
  b->n = 1;
  b->win.resize(1);
  WindowDesc& W = b->win[0];
  W = WindowDesc{};
  W.nc = W.ncf = ncf;
  W.nt = nt;
  W.rec = static_cast<int32_t>(rec.size());
  W.ncam = 1;
  W.no = 1;
  W.max_iterations = 1;
  std::vector<float4> poses(3 * ncf);
  for (int j = 0; j < ncf; ++j) {
    poses[3 * j] = make_float4(1, 0, 0, 0);
    poses[3 * j + 1] = make_float4(0, 1, 0, 0);
    poses[3 * j + 2] = make_float4(0, 0, 1, 0);
  }
  Batch& d = b->dev;
  d.n_windows = 1;
  int rc = SBA_OK;
  WindowState st0{};
  st0.lambda = 0.f;
  std::vector<WindowState> stv(1, st0);
  if ((rc = upload(h, b, b->win, const_cast<WindowDesc**>(&d.win))) != SBA_OK ||
      (rc = upload(h, b, stv, &d.st)) != SBA_OK || (rc = upload(h, b, rec, &d.sred)) != SBA_OK ||
      (rc = upload(h, b, poses, &d.poses0)) != SBA_OK || (rc = upload(h, b, poses, &d.poses1)) != SBA_OK ||
      (rc = alloc(h, b, static_cast<size_t>(n), &d.dc)) != SBA_OK || (rc = alloc(h, b, static_cast<size_t>(n), &d.dc64)) != SBA_OK ||
      (rc = alloc(h, b, 1, &d.A)) != SBA_OK) {
    free_batch(b);
    return rc;
  }
  if (launch_factor(d, nullptr, h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
    free_batch(b);
    return fail(h, SBA_E_DEVICE, "factor launch failed");
  }
  WindowState st1{};
  if (hipMemcpy(&st1, d.st, sizeof(st1), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(x, d.dc64, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) {
    free_batch(b);
    return fail(h, SBA_E_DEVICE, "factor readback failed");
  }
  free_batch(b);
  return st1.chol_ok ? SBA_OK : SBA_E_INFEASIBLE;
}

int sba_debug_mfma_layout(sba_handle* h, const float* A, const float* B, float* D) {
  if (!h || !A || !B || !D) return fail(h, SBA_E_INVALID, "bad debug_mfma_layout arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  float *a = nullptr, *b = nullptr, *d = nullptr;
  HIP_TRY(hipMalloc(&a, 64 * 4));
  HIP_TRY(hipMalloc(&b, 64 * 4));
  HIP_TRY(hipMalloc(&d, 256 * 4));
  HIP_TRY(hipMemcpy(a, A, 64 * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b, B, 64 * 4, hipMemcpyHostToDevice));
  HIP_TRY(launch_mfma_layout(a, b, d, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(D, d, 256 * 4, hipMemcpyDeviceToHost));
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(d);
  return SBA_OK;
}

int sba_debug_copy(sba_handle* h, int64_t bytes, int32_t repeats, float* ms) {
  if (!h || bytes <= 0 || repeats <= 0) return fail(h, SBA_E_INVALID, "bad debug_copy arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  const int64_t n = bytes / 16;
  void *a = nullptr, *c = nullptr;
  HIP_TRY(hipMalloc(&a, n * 16));
  HIP_TRY(hipMalloc(&c, n * 16));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipMemsetAsync(a, 0, n * 16, h->stream));
  HIP_TRY(hipEventRecord(e0, h->stream));
  for (int r = 0; r < repeats; ++r) HIP_TRY(launch_copy_calibrate(static_cast<float4*>(c), static_cast<const float4*>(a), n, h->stream));
  HIP_TRY(hipEventRecord(e1, h->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t / repeats;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(c);
  return SBA_OK;
}

}  // extern "C"

namespace {

// Workspace of the implicit-differentiation path of window 0 (FP64, one allocation).
struct AdjWork {
  int n6, nx, no;
  double *x, *r, *z, *p, *y, *ec, *zc, *acc, *gd, *sc, *Lst, *Linv, *dxy, *dinfo, *spts, *sposes;
  int32_t* done;
};

int adj_workspace(sba_handle* h, sba_batch* b, AdjWork& A) {
  const WindowDesc& W = b->win[0];
  A.n6 = 6 * W.ncf;
  A.nx = A.n6 + 3 * W.npf;
  A.no = W.no;
  const int64_t npk = static_cast<int64_t>(A.n6) * (A.n6 + 1) / 2;
  // x r z p y ec zc acc g (nx each), sc (8), done (2), factor (npk + n6), its inverse (n6 x n6), dxy (2 no),
  // dinfo (3 no), FP64 points (3 np) and key frames (12 nc)
  const size_t nx = static_cast<size_t>(A.nx);
  const size_t nd = 9 * nx + 8 + 2 + static_cast<size_t>(npk + A.n6) + static_cast<size_t>(A.n6) * A.n6 +
                    5 * static_cast<size_t>(A.no) + 3 * static_cast<size_t>(W.np) + 12 * static_cast<size_t>(W.nc);
  if (!b->adj) {
    HIP_TRY(hipMalloc(&b->adj, nd * sizeof(double)));
    b->allocs.push_back(b->adj);
  }
  A.x = b->adj;
  A.r = A.x + nx;
  A.z = A.r + nx;
  A.p = A.z + nx;
  A.y = A.p + nx;
  A.ec = A.y + nx;
  A.zc = A.ec + nx;
  A.acc = A.zc + nx;
  A.gd = A.acc + nx;
  A.sc = A.gd + nx;
  A.done = reinterpret_cast<int32_t*>(A.sc + 8);
  A.Lst = A.sc + 10;
  A.Linv = A.Lst + npk + A.n6;
  A.dxy = A.Linv + static_cast<size_t>(A.n6) * A.n6;
  A.dinfo = A.dxy + 2 * static_cast<size_t>(A.no);
  A.spts = A.dinfo + 3 * static_cast<size_t>(A.no);
  A.sposes = A.spts + 3 * static_cast<size_t>(W.np);
  return SBA_OK;
}

// Factors the Gauss-Newton system of the solution of window 0 (lambda = 0) once and copies its state to FP64.
// The batch's LM state is left as it was.
int adj_prepare(sba_handle* h, sba_batch* b, AdjWork& A) {
  if (b->win[0].mapped)
    return fail(h, SBA_E_CAPACITY, "implicit differentiation supports windows of at most 40 free key frames");
  int rc = adj_workspace(h, b, A);
  if (rc != SBA_OK || b->adj_ready) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  Batch d = b->dev;
  hipStream_t s = h->stream;
  WindowState st0;
  HIP_TRY(hipMemcpy(&st0, d.st, sizeof(st0), hipMemcpyDeviceToHost));
  WindowState st1 = st0;
  st1.lambda = 0.f;
  st1.done = 0;
  HIP_TRY(hipMemcpy(d.st, &st1, sizeof(st1), hipMemcpyHostToDevice));
  std::vector<WindowState> others;
  if (b->n > 1) {  // only window 0 runs: mark the others done for these launches
    others.resize(b->n - 1);
    HIP_TRY(hipMemcpy(others.data(), d.st + 1, sizeof(WindowState) * (b->n - 1), hipMemcpyDeviceToHost));
    std::vector<WindowState> tmp = others;
    for (auto& t : tmp) t.done = 1;
    HIP_TRY(hipMemcpy(d.st + 1, tmp.data(), sizeof(WindowState) * (b->n - 1), hipMemcpyHostToDevice));
  }
  WindowState chk{};
  if (launch_adj_state(d, A.spts, A.sposes, nullptr, 0, s) != hipSuccess || launch_schur(d, s) != hipSuccess ||
      launch_reduce(d, s) != hipSuccess || launch_factor(d, nullptr, s, A.Lst) != hipSuccess ||
      launch_adj_linv(A.Lst, A.n6, A.Linv, s) != hipSuccess) {
    rc = fail(h, SBA_E_DEVICE, "adjoint factor launch failed");
  } else if (hipMemcpyAsync(&chk, d.st, sizeof(chk), hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess) {
    rc = fail(h, SBA_E_DEVICE, "adjoint state read failed");
  } else if (!chk.chol_ok) {
    rc = fail(h, SBA_E_INFEASIBLE, "the Gauss-Newton system of the solution is not positive definite");
  }
  HIP_TRY(hipMemcpy(d.st, &st0, sizeof(st0), hipMemcpyHostToDevice));
  if (!others.empty())
    HIP_TRY(hipMemcpy(d.st + 1, others.data(), sizeof(WindowState) * others.size(), hipMemcpyHostToDevice));
  if (rc == SBA_OK) b->adj_ready = true;
  return rc;
}

// x = H^-1 gd by preconditioned conjugate gradients; the iteration count lands in sc[3]. The iterations are
// enqueued in blocks of kCgBlock and the host reads the device's convergence flag after each block, so a solve
// that converges early does not enqueue the remaining iterations (which would only idle on the flag).
constexpr int kCgBlock = 8;
hipError_t adj_pcg(const Batch& d, const AdjWork& A, int max_cg, double tol, hipStream_t s) {
  auto vec = [&](int op) {
    return launch_adj_vec(d, A.x, A.r, A.z, A.p, A.y, A.ec, A.zc, A.acc, A.gd, A.sc, A.done, tol, A.nx, A.n6, op, s);
  };
  auto precond = [&]() {
    hipError_t e = launch_adj_obs(d, 1, A.spts, A.sposes, A.r, nullptr, A.ec, nullptr, nullptr, nullptr, nullptr, A.done, s);
    if (e == hipSuccess) e = launch_adj_chol(A.Linv, A.n6, A.ec, A.zc, A.done, s);
    if (e == hipSuccess)
      e = launch_adj_obs(d, 2, A.spts, A.sposes, nullptr, A.acc, nullptr, A.zc, nullptr, nullptr, nullptr, A.done, s);
    if (e == hipSuccess) e = vec(1);
    return e;
  };
  hipError_t e = hipMemsetAsync(A.done, 0, sizeof(int32_t), s);
  if (e == hipSuccess) e = vec(0);
  if (e == hipSuccess) e = precond();
  if (e == hipSuccess) e = vec(2);
  int32_t done = 0;
  for (int it = 0; it < max_cg && e == hipSuccess && !done; ++it) {
    e = launch_adj_obs(d, 0, A.spts, A.sposes, A.p, A.y, nullptr, nullptr, nullptr, nullptr, nullptr, A.done, s);
    if (e == hipSuccess) e = vec(3);
    if (e == hipSuccess) e = precond();
    if (e == hipSuccess) e = vec(4);
    if (e == hipSuccess && (it + 1) % kCgBlock == 0 && it + 1 < max_cg) {
      e = hipMemcpyAsync(&done, A.done, sizeof(done), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
  }
  return e;
}

// z = M^-1 gd with the factored Gauss-Newton preconditioner M = 2 sum rho' J^T L J (one application, no
// conjugate gradients): a Gauss-Newton step.
hipError_t adj_gn_step(const Batch& d, const AdjWork& A, hipStream_t s) {
  hipError_t e = hipMemsetAsync(A.done, 0, sizeof(int32_t), s);
  if (e == hipSuccess)
    e = launch_adj_vec(d, A.x, A.r, A.z, A.p, A.y, A.ec, A.zc, A.acc, A.gd, A.sc, A.done, 0.0, A.nx, A.n6, 0, s);
  if (e == hipSuccess)
    e = launch_adj_obs(d, 1, A.spts, A.sposes, A.r, nullptr, A.ec, nullptr, nullptr, nullptr, nullptr, A.done, s);
  if (e == hipSuccess) e = launch_adj_chol(A.Linv, A.n6, A.ec, A.zc, A.done, s);
  if (e == hipSuccess)
    e = launch_adj_obs(d, 2, A.spts, A.sposes, nullptr, A.acc, nullptr, A.zc, nullptr, nullptr, nullptr, A.done, s);
  if (e == hipSuccess)
    e = launch_adj_vec(d, A.x, A.r, A.z, A.p, A.y, A.ec, A.zc, A.acc, A.gd, A.sc, A.done, 0.0, A.nx, A.n6, 1, s);
  return e;
}

// gd = gradient of the cost at the FP64 state.
hipError_t adj_grad(const Batch& d, const AdjWork& A, hipStream_t s) {
  hipError_t e = hipMemsetAsync(A.gd, 0, sizeof(double) * A.nx, s);
  if (e == hipSuccess)
    e = launch_adj_obs(d, 4, A.spts, A.sposes, nullptr, A.gd, nullptr, nullptr, nullptr, nullptr, nullptr, A.done, s);
  return e;
}

}  // namespace

extern "C" {

int sba_batch_refine(sba_handle* h, sba_batch* b, int32_t gauss_newton_steps, int32_t newton_steps, int32_t max_cg,
                     double tol, double* points, double* rig_from_world, double* grad_norm) {
  if (!h || !b || gauss_newton_steps < 0 || newton_steps < 0 || max_cg < 1)
    return fail(h, SBA_E_INVALID, "bad sba_batch_refine arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  AdjWork A{};
  int rc = adj_prepare(h, b, A);
  if (rc != SBA_OK) return rc;
  const Batch d = b->dev;
  const WindowDesc& W = b->win[0];
  hipStream_t s = h->stream;
  // Gauss-Newton steps x <- x - M^-1 grad: the factored system is positive definite, so these descend safely
  // from where the FP32 solve stopped, into the region where the full Hessian is positive definite too
  for (int k = 0; k < gauss_newton_steps; ++k) {
    HIP_TRY(adj_grad(d, A, s));
    HIP_TRY(adj_gn_step(d, A, s));
    HIP_TRY(launch_adj_state(d, A.spts, A.sposes, A.z, 1, s));
  }
  // Newton steps x <- x - H^-1 grad, full FP64 Hessian, the stored factor as preconditioner
  for (int k = 0; k < newton_steps; ++k) {
    HIP_TRY(adj_grad(d, A, s));
    HIP_TRY(adj_pcg(d, A, max_cg, tol, s));
    HIP_TRY(launch_adj_state(d, A.spts, A.sposes, A.x, 1, s));
  }
  if (grad_norm) {
    HIP_TRY(adj_grad(d, A, s));
    std::vector<double> g(A.nx);
    HIP_TRY(hipMemcpyAsync(g.data(), A.gd, sizeof(double) * A.nx, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double n2 = 0.0;
    for (double v : g) n2 += v * v;
    *grad_norm = std::sqrt(n2);
  }
  if (points || rig_from_world) {
    std::vector<double> P(3 * static_cast<size_t>(W.np)), T(12 * static_cast<size_t>(W.nc));
    HIP_TRY(hipMemcpyAsync(P.data(), A.spts, sizeof(double) * P.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(T.data(), A.sposes, sizeof(double) * T.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (points) std::copy(P.begin(), P.end(), points);
    if (rig_from_world)
      for (int j = 0; j < W.nc; ++j) {
        double* M = rig_from_world + 16 * static_cast<size_t>(j);  // column-major 4x4
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 4; ++c) M[4 * c + r] = T[12 * static_cast<size_t>(j) + 4 * r + c];
        M[3] = M[7] = M[11] = 0.0;
        M[15] = 1.0;
      }
  }
  return SBA_OK;
}

int sba_autotune_chunks(sba_handle* h, const sba_problem_view* problems, int32_t n, const int32_t* candidates,
                        int32_t n_candidates, int32_t repeats, int32_t* best_chunk_pairs, float* ms_per_candidate) {
  if (!h || !problems || n <= 0 || !best_chunk_pairs || repeats < 1 || n_candidates < 0 || (n_candidates > 0 && !candidates))
    return fail(h, SBA_E_INVALID, "bad sba_autotune_chunks arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  int64_t total_pairs = 0;
  {
    // the default candidates bracket the heuristic: x1/4 .. x4
    sba_batch* probe = nullptr;
    const int32_t saved = h->opt.chunk_pairs;
    h->opt.chunk_pairs = 0;
    const int rc = sba_batch_create(h, problems, n, &probe);
    h->opt.chunk_pairs = saved;
    if (rc != SBA_OK) return rc;
    total_pairs = probe->total_pairs;
    sba_batch_destroy(h, probe);
  }
  std::vector<int32_t> cand(candidates, candidates + n_candidates);
  if (cand.empty()) {
    const int64_t base = default_chunk_pairs(total_pairs);
    for (int64_t c : {base / 4, base / 2, base, 2 * base, 4 * base})
      cand.push_back(static_cast<int32_t>(std::max<int64_t>(256, c)));
  }
  const int32_t saved = h->opt.chunk_pairs;
  float best = 0.f;
  int32_t best_c = cand[0];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = SBA_OK;
  for (size_t k = 0; k < cand.size() && rc == SBA_OK; ++k) {
    h->opt.chunk_pairs = cand[k];
    sba_batch* b = nullptr;
    rc = sba_batch_create(h, problems, n, &b);
    if (rc != SBA_OK) break;
    float ms = 0.f;
    if ((rc = sba_batch_run(h, b)) == SBA_OK && hipEventRecord(e0, h->stream) == hipSuccess) {
      for (int r = 0; r < repeats && rc == SBA_OK; ++r) rc = sba_batch_run(h, b);
      if (rc == SBA_OK && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(h, SBA_E_DEVICE, "autotune timing failed");
    }
    sba_batch_destroy(h, b);
    ms /= static_cast<float>(repeats);
    if (ms_per_candidate && k < static_cast<size_t>(n_candidates)) ms_per_candidate[k] = ms;
    if (rc == SBA_OK && (k == 0 || ms < best)) {
      best = ms;
      best_c = cand[k];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->opt.chunk_pairs = saved;
  if (rc == SBA_OK) *best_chunk_pairs = best_c;
  return rc;
}

int sba_autotune_table(sba_handle* h, const char* path, const sba_problem_view* problems, int32_t n, int32_t repeats,
                       int32_t* chunk_pairs, int32_t* bchunk_pairs, int32_t* from_table) {
  if (!h || !path || !problems || n <= 0 || !chunk_pairs || !bchunk_pairs || repeats < 1)
    return fail(h, SBA_E_INVALID, "bad sba_autotune_table arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  // the batch's shape class: device architecture, loss, window count, total observations and pairs to two
  // significant digits, largest free key-frame count
  hipDeviceProp_t prop{};
  HIP_TRY(hipGetDeviceProperties(&prop, h->opt.device));
  int64_t no = 0, maxf = 0;
  for (int w = 0; w < n; ++w) {
    no += problems[w].num_observations;
    maxf = std::max<int64_t>(maxf, problems[w].num_poses - problems[w].num_fixed_key_frames);
  }
  auto two_digits = [](int64_t v) {
    int64_t scale = 1;
    while (v >= 100) {
      v /= 10;
      scale *= 10;
    }
    return v * scale;
  };
  char key[512];
  std::snprintf(key, sizeof(key), "%s loss%d windows%d obs%lld maxfree%lld", prop.gcnArchName, h->opt.loss, n,
                static_cast<long long>(two_digits(no)), static_cast<long long>(maxf));
  for (char* c = key; *c; ++c)
    if (*c == ':' ) *c = '_';
  if (from_table) *from_table = 0;
  if (FILE* f = std::fopen(path, "r")) {
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
      char k[256];
      int c = 0, bc = 0;
      // "<key fields...> | <chunk_pairs> <bchunk_pairs>"
      const char* bar = std::strchr(line, '|');
      if (!bar || static_cast<size_t>(bar - line) >= sizeof(k)) continue;
      std::memcpy(k, line, bar - line);
      k[bar - line] = 0;
      while (bar - line > 0 && k[std::strlen(k) - 1] == ' ') k[std::strlen(k) - 1] = 0;
      if (std::strcmp(k, key) == 0 && std::sscanf(bar + 1, "%d %d", &c, &bc) == 2 && c > 0 && bc > 0) {
        *chunk_pairs = c;
        *bchunk_pairs = bc;
        if (from_table) *from_table = 1;
        std::fclose(f);
        return SBA_OK;
      }
    }
    std::fclose(f);
  }
  // not in the table: tune the Schur chunks, then the back-substitution chunks at the chosen Schur size
  int32_t best = 0;
  int rc = sba_autotune_chunks(h, problems, n, nullptr, 0, repeats, &best, nullptr);
  if (rc != SBA_OK) return rc;
  const int32_t saved_c = h->opt.chunk_pairs, saved_b = h->opt.bchunk_pairs;
  h->opt.chunk_pairs = best;
  int32_t best_b = best;
  float best_ms = 0.f;
  for (int32_t m : {1, 2, 4, 8}) {
    h->opt.bchunk_pairs = best * m;
    sba_batch* b = nullptr;
    if ((rc = sba_batch_create(h, problems, n, &b)) != SBA_OK) break;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess && sba_batch_run(h, b) == SBA_OK &&
        hipEventRecord(e0, h->stream) == hipSuccess) {
      for (int r = 0; r < repeats; ++r) (void)sba_batch_run(h, b);
      if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = fail(h, SBA_E_DEVICE, "autotune timing failed");
    } else {
      rc = fail(h, SBA_E_DEVICE, "autotune launch failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    sba_batch_destroy(h, b);
    if (rc != SBA_OK) break;
    if (m == 1 || ms < best_ms) {
      best_ms = ms;
      best_b = best * m;
    }
  }
  h->opt.chunk_pairs = saved_c;
  h->opt.bchunk_pairs = saved_b;
  if (rc != SBA_OK) return rc;
  *chunk_pairs = best;
  *bchunk_pairs = best_b;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%s | %d %d %.4f\n", key, best, best_b, best_ms / repeats);
    std::fclose(f);
  }
  return SBA_OK;
}

int sba_batch_adjoint(sba_handle* h, sba_batch* b, const double* g, double* dxy, double* dinfo, int32_t max_cg,
                      double tol, int32_t* cg_iterations, double* a_out) {
  if (!h || !b || !g || !dxy || !dinfo || max_cg < 1) return fail(h, SBA_E_INVALID, "bad sba_batch_adjoint arguments");
  HIP_TRY(hipSetDevice(h->opt.device));
  AdjWork A{};
  int rc = adj_prepare(h, b, A);
  if (rc != SBA_OK) return rc;
  const Batch d = b->dev;
  hipStream_t s = h->stream;
  HIP_TRY(hipMemcpyAsync(A.gd, g, sizeof(double) * A.nx, hipMemcpyHostToDevice, s));
  HIP_TRY(adj_pcg(d, A, max_cg, tol, s));
  // per-observation gradients from the adjoint
  HIP_TRY(launch_adj_obs(d, 3, A.spts, A.sposes, A.x, nullptr, nullptr, nullptr, A.dxy, A.dinfo, b->obs_orig, A.done, s));
  double scal[8];
  HIP_TRY(hipMemcpyAsync(scal, A.sc, sizeof(scal), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(dxy, A.dxy, sizeof(double) * 2 * A.no, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(dinfo, A.dinfo, sizeof(double) * 3 * A.no, hipMemcpyDeviceToHost, s));
  if (a_out) HIP_TRY(hipMemcpyAsync(a_out, A.x, sizeof(double) * A.nx, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (cg_iterations) *cg_iterations = static_cast<int32_t>(scal[3]);
  return SBA_OK;
}

int sba_solve_batch(sba_handle* h, const sba_problem_view* problems, int32_t n, sba_result* results) {
  sba_batch* b = nullptr;
  int rc = sba_batch_create(h, problems, n, &b);
  if (rc != SBA_OK) return rc;
  rc = sba_batch_run(h, b);
  if (rc == SBA_OK) rc = sba_batch_read(h, b, results);
  sba_batch_destroy(h, b);
  return rc;
}

int sba_solve(sba_handle* h, const sba_problem_view* problem, sba_result* result) {
  return solve_one(h, problem, result, [h](sba_batch*, size_t bytes, char** base) {
    return workspace_memory(h, bytes, base);
  });
}

int sba_solve_noalloc(sba_handle* h, const sba_problem_view* problem, sba_result* result) {
  return solve_one(h, problem, result, [h](sba_batch*, size_t bytes, char** base) -> int {
    if (!h->scratch) return fail(h, SBA_E_DEVICE, "the scratch array of sba_solve_noalloc is not set up");
    if (bytes > h->scratch_bytes)
      return fail(h, SBA_E_CAPACITY, "window needs " + std::to_string(bytes) + " bytes of device memory; "
                  "sba_solve_noalloc holds at most " + std::to_string(h->scratch_bytes) + " (use sba_solve)");
    *base = h->scratch;
    return SBA_OK;
  });
}

}  // extern "C"
