// Landmark winsorizing between two solves of a window (cuVSLAM: pipelines/winsorizer.cpp:23-73, called between
// the two bundler runs of service_sba.h:317-339), on the solved state of a device-resident batch:
//   k_wins_err     one thread per point of the window: the FP64 squared reprojection error (normalized
//                  coordinates) of each observation made from the chosen key frame, -1 for every other
//                  observation and for points inside the near plane (IsDepthInFront, near_plane.h:38), and the
//                  number of such observations;
//   k_wins_select  one workgroup: with at least 50 of them, the int(5 n / 100) largest errors by a radix select
//                  over the error's bit pattern (8 passes of 8 bits; non-negative doubles order as their bits),
//                  ties at the cut taken by decreasing point index; those above 0.0025^2 mark their points;
//   k_wins_apply   one thread per point: the observations of marked points get zero information, so a later
//                  solve excludes them as the reference's second run excludes reset landmarks, and the window's
//                  current state becomes its initial state, so that the later solve continues from this one.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "sba_device.h"

using namespace sba_dev;

namespace {

constexpr int kSelThreads = 1024;
constexpr int kMaxTies = 1024;          // observations sharing the cut error
constexpr int kSafe = 50;               // winsorizer.cpp:29
constexpr float kPercent = 5.f;         // winsorizer.cpp:27
constexpr float kMinSize = 0.0025f;     // winsorizer.cpp:28
constexpr float kHitherW = 0.1f;        // epipolar/near_plane.h:33-38

}  // namespace

__global__ __launch_bounds__(256) void k_wins_err(Batch b, int w, int kf, double* err, int32_t* pt_of,
                                                  int32_t* count) {
  const WindowDesc W = b.win[w];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.np) return;
  const int cur = b.st[w].cur;
  const float4* pts = cur ? b.pts1 : b.pts0;
  const float4* P = (cur ? b.poses1 : b.poses0) + 3 * (W.pose_off + kf);
  const float4* cams = b.cams + 3 * W.cam_off;
  const int gi = W.point_off + i;
  const float4 x = pts[gi];
  const float4 p0 = P[0], p1 = P[1], p2 = P[2];
  // rig point of the key frame, FP64 from the FP32 state
  const double rx = static_cast<double>(p0.x) * x.x + static_cast<double>(p0.y) * x.y +
                    static_cast<double>(p0.z) * x.z + static_cast<double>(p0.w);
  const double ry = static_cast<double>(p1.x) * x.x + static_cast<double>(p1.y) * x.y +
                    static_cast<double>(p1.z) * x.z + static_cast<double>(p1.w);
  const double rz = static_cast<double>(p2.x) * x.x + static_cast<double>(p2.y) * x.y +
                    static_cast<double>(p2.z) * x.z + static_cast<double>(p2.w);
  int n = 0;
  for (int q = b.point_pair[gi]; q < b.point_pair[gi + 1]; ++q) {
    const bool seen = b.pair_pose[q] == kf;
    for (int o = b.pair_rec[q].x, oe = b.pair_rec[q].y; o < oe; ++o) {
      double e = -1.0;
      if (seen) {
        const float4 inf = b.obs_info[o];
        const float4* C = cams + 3 * __float_as_int(inf.w);
        const float4 c0 = C[0], c1 = C[1], c2 = C[2];
        const double cx = c0.x * rx + c0.y * ry + c0.z * rz + c0.w;
        const double cy = c1.x * rx + c1.y * ry + c1.z * rz + c1.w;
        const double cz = c2.x * rx + c2.y * ry + c2.z * rz + c2.w;
        if (cz > static_cast<double>(kHitherW)) {
          const float2 xy = b.obs_xy[o];
          const double u = cx / cz - xy.x, v = cy / cz - xy.y;
          e = u * u + v * v;
          ++n;
        }
      }
      err[o - W.obs_off] = e;
      pt_of[o - W.obs_off] = i;
    }
  }
  if (n) atomicAdd(count, n);
}

// out: [0] points marked, [1] candidate observations, [2] observations cut, [3] 1 when more than kMaxTies
// observations share the cut error (nothing is marked then).
__global__ __launch_bounds__(kSelThreads) void k_wins_select(Batch b, int w, const double* err, const int32_t* pt_of,
                                                            const int32_t* count, uint8_t* mark, int32_t* out) {
  __shared__ unsigned int hist[256];
  __shared__ unsigned long long s_prefix;
  __shared__ int s_need, s_nties, s_marked;
  __shared__ int ties[kMaxTies];
  const WindowDesc W = b.win[w];
  const int tid = threadIdx.x;
  const int n = *count;
  const int k = static_cast<int>(kPercent * static_cast<float>(n) / 100.f);
  if (n < kSafe || k <= 0) {
    if (tid == 0) {
      out[0] = 0;
      out[1] = n;
      out[2] = 0;
      out[3] = 0;
    }
    return;
  }
  const double min2 = static_cast<double>(kMinSize * kMinSize);
  unsigned long long prefix = 0, sel = 0;
  int need = k;
  for (int pass = 7; pass >= 0; --pass) {
    const int sh = 8 * pass;
    for (int d = tid; d < 256; d += kSelThreads) hist[d] = 0;
    __syncthreads();
    for (int o = tid; o < W.no; o += kSelThreads) {
      const double e = err[o];
      if (e >= 0.0) {
        const unsigned long long key = static_cast<unsigned long long>(__double_as_longlong(e));
        if ((key & sel) == prefix) atomicAdd(&hist[(key >> sh) & 255ull], 1u);
      }
    }
    __syncthreads();
    if (tid == 0) {
      int cum = 0, d = 255;
      for (; d > 0; --d) {
        if (cum + static_cast<int>(hist[d]) >= need) break;
        cum += static_cast<int>(hist[d]);
      }
      s_need = need - cum;
      s_prefix = prefix | (static_cast<unsigned long long>(d) << sh);
    }
    __syncthreads();
    need = s_need;
    prefix = s_prefix;
    sel |= 0xFFull << sh;
    __syncthreads();
  }
  // prefix is the k-th largest error's bit pattern; `need` observations with exactly that error are cut
  if (tid == 0) {
    s_nties = 0;
    s_marked = 0;
  }
  __syncthreads();
  for (int o = tid; o < W.no; o += kSelThreads) {
    const double e = err[o];
    if (e < 0.0) continue;
    const unsigned long long key = static_cast<unsigned long long>(__double_as_longlong(e));
    if (key > prefix) {
      if (e > min2) mark[pt_of[o]] = 1;
    } else if (key == prefix) {
      const int slot = atomicAdd(&s_nties, 1);
      if (slot < kMaxTies) ties[slot] = o;
    }
  }
  __syncthreads();
  if (tid == 0) {
    const int nt = s_nties;
    if (nt <= kMaxTies) {
      // decreasing point index, then decreasing observation index
      for (int a = 1; a < nt; ++a) {
        const int v = ties[a];
        int c = a - 1;
        while (c >= 0 && (pt_of[ties[c]] < pt_of[v] || (pt_of[ties[c]] == pt_of[v] && ties[c] < v))) {
          ties[c + 1] = ties[c];
          --c;
        }
        ties[c + 1] = v;
      }
      for (int j = 0; j < need && j < nt; ++j)
        if (err[ties[j]] > min2) mark[pt_of[ties[j]]] = 1;
    }
    out[1] = n;
    out[2] = k;
    out[3] = nt > kMaxTies ? 1 : 0;
  }
  __syncthreads();
  if (s_nties > kMaxTies) {  // leave nothing half-marked
    for (int i = tid; i < W.np; i += kSelThreads) mark[i] = 0;
    if (tid == 0) out[0] = 0;
    return;
  }
  int m = 0;
  for (int i = tid; i < W.np; i += kSelThreads) m += mark[i] ? 1 : 0;
  if (m) atomicAdd(&s_marked, m);
  __syncthreads();
  if (tid == 0) out[0] = s_marked;
}

__global__ __launch_bounds__(256) void k_wins_apply(Batch b, int w, const uint8_t* mark) {
  const WindowDesc W = b.win[w];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int cur = b.st[w].cur;
  if (i < W.np) {
    const int gi = W.point_off + i;
    const_cast<float4*>(b.pts_init)[gi] = (cur ? b.pts1 : b.pts0)[gi];
    if (mark[i]) {
      float4* info = const_cast<float4*>(b.obs_info);
      float4* pinfo = const_cast<float4*>(b.pair_info);
      for (int q = b.point_pair[gi]; q < b.point_pair[gi + 1]; ++q) {
        for (int o = b.pair_rec[q].x, oe = b.pair_rec[q].y; o < oe; ++o) info[o] = make_float4(0.f, 0.f, 0.f, info[o].w);
        if (pinfo != info) pinfo[q] = make_float4(0.f, 0.f, 0.f, pinfo[q].w);
      }
    }
  }
  if (i < 3 * W.nc) {
    const int g = 3 * W.pose_off + i;
    const_cast<float4*>(b.poses_init)[g] = (cur ? b.poses1 : b.poses0)[g];
  }
}

namespace sba_dev {

hipError_t launch_winsorize_select(const Batch& b, int w, int np, int kf, double* err, int32_t* pt_of,
                                   int32_t* count, uint8_t* mark, int32_t* out, hipStream_t s) {
  const int gx = (np + 255) / 256;
  if (gx > 0) hipLaunchKernelGGL(k_wins_err, dim3(gx), dim3(256), 0, s, b, w, kf, err, pt_of, count);
  hipLaunchKernelGGL(k_wins_select, dim3(1), dim3(kSelThreads), 0, s, b, w, err, pt_of, count, mark, out);
  return hipGetLastError();
}

hipError_t launch_winsorize_apply(const Batch& b, int w, int np, int nc, const uint8_t* mark, hipStream_t s) {
  const int ga = (std::max(np, 3 * nc) + 255) / 256;
  if (ga > 0) hipLaunchKernelGGL(k_wins_apply, dim3(ga), dim3(256), 0, s, b, w, mark);
  return hipGetLastError();
}

}  // namespace sba_dev
