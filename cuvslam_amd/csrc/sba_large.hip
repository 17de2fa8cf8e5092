// Reduced camera system of windows with more than kMaxFreePoses free key frames ("mapped" windows, e.g. a large
// multi-camera map with a reduced system of order ~1200).
//
// The Schur kernel stages at most kMaxFreePoses free key frames per chunk; a mapped window's chunks carry the
// list of the free key frames they stage (chunk_map, increasing). Here:
//   k_zero_big    clears the window's dense FP32 sum of Z Z^T and its U, e, diag U
//   k_scatter     adds each chunk record into them at the rows of its key frames (FP32 atomics: the order of
//                 the additions, and so the last bits of the sum, is not fixed for these windows)
//   k_lassemble   A = U + lambda diag(U) - sum Z Z^T in FP64 (dense, row-major lower triangle), rhs e
//   k_potrf / k_trsm / k_syrk   right-looking blocked Cholesky (Golub & Van Loan, Matrix Computations,
//                 Algorithm 4.2.10) with 64 x 64 blocks: factor of the diagonal block in one workgroup, the
//                 panel below it one block per workgroup, the trailing update one block pair per workgroup
//   k_ltrsv       L y = e, L^T x = y in one workgroup, 64-row blocks
//   k_lepilogue   the pose step, its max-norm and dc^T diag(U) dc, trial key frames
// A non-positive pivot clears the window's chol_ok; the later kernels then return, and the LM decision raises
// lambda (schur_complement_bundler_gpu.cpp:175-185), as for the small windows.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "sba_device.h"

using namespace sba_dev;

namespace {

constexpr int kLB = 64;        // block of the large factorization
constexpr int kLds = kLB + 1;  // padded LDS row

__device__ __forceinline__ int nblocks(int n) { return (n + kLB - 1) / kLB; }

struct Large {
  double* A;    // n x n, row-major; the lower triangle holds A, then L
  double* rhs;  // e, then y, then x
  double* du;   // diag(U)
  int n;
};

__device__ __forceinline__ Large large_of(const Batch& b, const WindowDesc& W) {
  const int n = 6 * W.ncf;
  double* A = b.A + W.a_off;
  return Large{A, A + static_cast<int64_t>(n) * n, A + static_cast<int64_t>(n) * n + n, n};
}

__device__ __forceinline__ double readlane_d(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

}  // namespace

__global__ __launch_bounds__(256) void k_zero_big(Batch b, int w) {
  const WindowDesc W = b.win[w];
  if (b.st[w].done) return;
  const int64_t n = 6 * W.ncf;
  float* big = b.big + W.big_off;
  float* ue = b.big_ue + W.ue_off;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t i = g; i < n * n / 4; i += stride) reinterpret_cast<float4*>(big)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t i = 4 * (n * n / 4) + g; i < n * n; i += stride) big[i] = 0.f;
  for (int64_t i = g; i < 33 * W.ncf; i += stride) ue[i] = 0.f;
}

// One workgroup per chunk of window w: record element (local row rl, local column cl), rl <= cl, goes to
// (6 map[rl / 6] + rl % 6, 6 map[cl / 6] + cl % 6), which is again in the upper triangle (map increasing).
__global__ __launch_bounds__(256) void k_scatter(Batch b, int w) {
  __shared__ int smap[kMaxFreePoses];
  const WindowDesc W = b.win[w];
  if (b.st[w].done) return;
  const Chunk ch = b.chunks[W.chunk_begin + blockIdx.x];
  const int nloc = ch.nloc, nt = ch.nt, n = 6 * W.ncf, ncf = W.ncf;
  for (int k = threadIdx.x; k < nloc; k += blockDim.x) smap[k] = b.chunk_map[ch.map_off + k];
  __syncthreads();
  const float* rec = b.part + ch.part_off;
  float* big = b.big + W.big_off;
  float* ue = b.big_ue + W.ue_off;
  const int ntu = nt * (nt + 1) / 2, nl6 = 6 * nloc;
  for (int idx = threadIdx.x; idx < 256 * ntu; idx += blockDim.x) {
    int T = idx >> 8, I = 0;
    while (T >= nt - I) {
      T -= nt - I;
      ++I;
    }
    const int J = I + T;
    const int e = idx & 255, v = e >> 6, lane = e & 63;
    const int rl = 16 * I + 4 * (lane >> 4) + v, cl = 16 * J + (lane & 15);
    if (rl >= nl6 || cl >= nl6 || rl > cl) continue;
    const int R = 6 * smap[rl / 6] + rl % 6, C = 6 * smap[cl / 6] + cl % 6;
    atomicAdd(big + static_cast<int64_t>(R) * n + C, rec[idx]);
  }
  const float* tail = rec + 256 * ntu;
  for (int idx = threadIdx.x; idx < 33 * nloc; idx += blockDim.x) {
    const float x = tail[idx];
    if (idx < 21 * nloc) {
      atomicAdd(ue + 21 * smap[idx / 21] + idx % 21, x);
    } else if (idx < 27 * nloc) {
      const int q = idx - 21 * nloc;
      atomicAdd(ue + 21 * ncf + 6 * smap[q / 6] + q % 6, x);
    } else {
      const int q = idx - 27 * nloc;
      atomicAdd(ue + 27 * ncf + 6 * smap[q / 6] + q % 6, x);
    }
  }
}

// A = U + lambda diag(U) - sum Z Z^T (lower triangle), rhs = e, du = diag(U); chol_ok set for the factor.
__global__ __launch_bounds__(256) void k_lassemble(Batch b, int w) {
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done) return;
  const Large L = large_of(b, W);
  const int n = L.n, ncf = W.ncf;
  const float* big = b.big + W.big_off;
  const float* U = b.big_ue + W.ue_off;
  const float* E = U + 21 * ncf;
  const float* D = E + 6 * ncf;
  const double lam = static_cast<double>(St->lambda);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < static_cast<int64_t>(n) * n;
       idx += stride) {
    const int R = static_cast<int>(idx / n), C = static_cast<int>(idx % n);
    if (C > R) continue;
    double v = -static_cast<double>(big[static_cast<int64_t>(C) * n + R]);
    if (R / 6 == C / 6) {
      const int kf = R / 6, rr = C - 6 * kf, cc = R - 6 * kf;
      v += static_cast<double>(U[21 * kf + 6 * rr - (rr * (rr - 1)) / 2 + (cc - rr)]);
    }
    if (R == C) v += lam * static_cast<double>(D[R]);
    L.A[idx] = v;
  }
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    L.rhs[i] = static_cast<double>(E[i]);
    L.du[i] = static_cast<double>(D[i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) St->chol_ok = 1;
}

// Cholesky of diagonal block k (one workgroup): column by column in LDS.
__global__ __launch_bounds__(256) void k_potrf(Batch b, int w, int k) {
  __shared__ double s[kLB * kLds];
  __shared__ double sinv;
  __shared__ int fail;
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  for (int idx = tid; idx < kLB * kLB; idx += blockDim.x) {
    const int r = idx / kLB, c = idx % kLB;
    double v = 0.0;
    if (r < kb && c < kb) {
      if (c <= r) v = L.A[static_cast<int64_t>(k0 + r) * n + k0 + c];
    } else if (r == c) {
      v = 1.0;  // identity padding beyond the matrix
    }
    s[r * kLds + c] = v;
  }
  if (tid == 0) fail = 0;
  for (int c = 0; c < kLB; ++c) {
    __syncthreads();
    if (tid == 0) {
      const double d = s[c * kLds + c];
      if (!(d > 0.0)) fail = 1;
      const double l = sqrt(fmax(d, 1e-300));
      s[c * kLds + c] = l;
      sinv = 1.0 / l;
    }
    __syncthreads();
    if (tid > c && tid < kLB) s[tid * kLds + c] *= sinv;
    __syncthreads();
    for (int r = c + 1 + (tid >> 6); r < kLB; r += 4)
      for (int j = c + 1 + (tid & 63); j <= r; j += 64) s[r * kLds + j] -= s[r * kLds + c] * s[j * kLds + c];
  }
  __syncthreads();
  for (int idx = tid; idx < kb * kb; idx += blockDim.x) {
    const int r = idx / kb, c = idx % kb;
    if (c <= r) L.A[static_cast<int64_t>(k0 + r) * n + k0 + c] = s[r * kLds + c];
  }
  if (tid == 0 && fail) St->chol_ok = 0;
}

// Panel below diagonal block k: L_ik = A_ik L_kk^-T, one block row per workgroup.
__global__ __launch_bounds__(256) void k_trsm(Batch b, int w, int k) {
  __shared__ double sL[kLB * kLds];
  __shared__ double sB[kLB * kLds];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  const int i0 = kLB * (k + 1 + blockIdx.x), ib = min(kLB, n - i0);
  for (int idx = tid; idx < kLB * kLB; idx += blockDim.x) {
    const int r = idx / kLB, c = idx % kLB;
    sL[r * kLds + c] = (r < kb && c <= r) ? L.A[static_cast<int64_t>(k0 + r) * n + k0 + c] : 0.0;
    sB[r * kLds + c] = (r < ib && c < kb) ? L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  if (tid < ib) {
    double* x = sB + tid * kLds;
    for (int c = 0; c < kb; ++c) {
      double a = x[c];
      for (int m = 0; m < c; ++m) a -= x[m] * sL[c * kLds + m];
      x[c] = a / sL[c * kLds + c];
    }
  }
  __syncthreads();
  for (int idx = tid; idx < ib * kb; idx += blockDim.x) {
    const int r = idx / kb, c = idx % kb;
    L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] = sB[r * kLds + c];
  }
}

// Trailing update after block k: A_ij -= L_ik L_jk^T for k < j <= i, one block pair per workgroup, each thread
// a 4 x 4 piece of the 64 x 64 block.
__global__ __launch_bounds__(256) void k_syrk(Batch b, int w, int k) {
  __shared__ double sA[kLB * kLds];
  __shared__ double sB[kLB * kLds];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  int ii = 0, t = blockIdx.x;
  while (t > ii) {
    t -= ii + 1;
    ++ii;
  }
  const int jj = t;  // 0 <= jj <= ii
  const int i0 = kLB * (k + 1 + ii), j0 = kLB * (k + 1 + jj);
  const int ib = min(kLB, n - i0), jb = min(kLB, n - j0);
  for (int idx = tid; idx < kLB * kLB; idx += blockDim.x) {
    const int r = idx / kLB, c = idx % kLB;
    sA[r * kLds + c] = (r < ib && c < kb) ? L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] : 0.0;
    sB[r * kLds + c] = (r < jb && c < kb) ? L.A[static_cast<int64_t>(j0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  const int r0 = 4 * (tid >> 4), c0 = 4 * (tid & 15);
  double acc[4][4] = {};
  for (int m = 0; m < kb; ++m) {
    double a[4], bb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = sA[(r0 + q) * kLds + m];
      bb[q] = sB[(c0 + q) * kLds + m];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * bb[q];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + p, c = c0 + q;
      if (r < ib && c < jb && (ii != jj || c <= r)) L.A[static_cast<int64_t>(i0 + r) * n + j0 + c] -= acc[p][q];
    }
}

// L y = e, then L^T x = y, in one workgroup of 1024 threads: per 64-row block, the off-diagonal products are
// summed by 16 thread slices, then wavefront 0 solves the diagonal block with register broadcasts.
__global__ __launch_bounds__(1024) void k_ltrsv(Batch b, int w) {
  extern __shared__ double sy[];  // n: y, then x
  __shared__ double part[16 * kLB];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, nb = nblocks(n), tid = threadIdx.x, lane = tid & 63, slice = tid >> 6;
  for (int i = tid; i < n; i += blockDim.x) sy[i] = L.rhs[i];
  __syncthreads();
  for (int blk = 0; blk < nb; ++blk) {
    const int b0 = kLB * blk, bb = min(kLB, n - b0);
    double s = 0.0;
    if (lane < bb)
      for (int c = slice; c < b0; c += 16) s += L.A[static_cast<int64_t>(b0 + lane) * n + c] * sy[c];
    part[slice * kLB + lane] = s;
    __syncthreads();
    if (slice == 0) {
      double v = 0.0;
      if (lane < bb) {
        v = sy[b0 + lane];
        for (int q = 0; q < 16; ++q) v -= part[q * kLB + lane];
      }
      for (int c = 0; c < bb; ++c) {
        const double yc = readlane_d(v, c) / L.A[static_cast<int64_t>(b0 + c) * n + b0 + c];
        if (lane == c) v = yc;
        else if (lane > c && lane < bb) v -= L.A[static_cast<int64_t>(b0 + lane) * n + b0 + c] * yc;
      }
      if (lane < bb) sy[b0 + lane] = v;
    }
    __syncthreads();
  }
  for (int blk = nb - 1; blk >= 0; --blk) {
    const int b0 = kLB * blk, bb = min(kLB, n - b0);
    double s = 0.0;
    if (lane < bb)
      for (int r = b0 + kLB + slice; r < n; r += 16) s += L.A[static_cast<int64_t>(r) * n + b0 + lane] * sy[r];
    part[slice * kLB + lane] = s;
    __syncthreads();
    if (slice == 0) {
      double v = 0.0;
      if (lane < bb) {
        v = sy[b0 + lane];
        for (int q = 0; q < 16; ++q) v -= part[q * kLB + lane];
      }
      for (int c = bb - 1; c >= 0; --c) {
        const double xc = readlane_d(v, c) / L.A[static_cast<int64_t>(b0 + c) * n + b0 + c];
        if (lane == c) v = xc;
        else if (lane < c) v -= L.A[static_cast<int64_t>(b0 + c) * n + b0 + lane] * xc;
      }
      if (lane < bb) sy[b0 + lane] = v;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) L.rhs[i] = sy[i];
}

// The pose step, max |dc|, dc^T diag(U) dc and the trial key frames of window w.
__global__ __launch_bounds__(256) void k_lepilogue(Batch b, int w) {
  __shared__ double red[8];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, tid = threadIdx.x, lane = tid & 63;
  double sc = 0.0;
  float mx = 0.f;
  for (int i = tid; i < n; i += blockDim.x) {
    const double x = L.rhs[i];
    b.dc[W.dc_off + i] = static_cast<float>(x);
    b.dc64[W.dc_off + i] = x;
    sc += x * x * L.du[i];
    mx = fmaxf(mx, fabsf(static_cast<float>(x)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sc += __shfl_xor(sc, o);
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  if (lane == 0) {
    red[tid >> 6] = sc;
    red[4 + (tid >> 6)] = static_cast<double>(mx);
  }
  __syncthreads();
  if (tid == 0) {
    St->scal_pose = red[0] + red[1] + red[2] + red[3];
    St->maxdc = static_cast<float>(fmax(fmax(red[4], red[5]), fmax(red[6], red[7])));
  }
  write_trial_poses(b, W, St->cur, L.rhs, tid, blockDim.x);
}

namespace sba_dev {

hipError_t launch_large_zero(const Batch& b, int w, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_big, dim3(256), dim3(256), 0, s, b, w);
  return hipGetLastError();
}

hipError_t launch_large_scatter(const Batch& b, int w, int nchunks, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(nchunks), dim3(256), 0, s, b, w);
  return hipGetLastError();
}

// The factor and solve of window w (n = 6 free key frames): assemble, blocked Cholesky, triangular solves,
// epilogue. 3 nb launches for nb 64-row blocks.
hipError_t launch_large_factor(const Batch& b, int w, int n, hipStream_t s) {
  const int nb = (n + kLB - 1) / kLB;
  hipLaunchKernelGGL(k_lassemble, dim3(512), dim3(256), 0, s, b, w);
  hipError_t e = hipGetLastError();
  for (int k = 0; k < nb && e == hipSuccess; ++k) {
    hipLaunchKernelGGL(k_potrf, dim3(1), dim3(256), 0, s, b, w, k);
    const int m = nb - k - 1;
    if (m > 0) {
      hipLaunchKernelGGL(k_trsm, dim3(m), dim3(256), 0, s, b, w, k);
      hipLaunchKernelGGL(k_syrk, dim3(m * (m + 1) / 2), dim3(256), 0, s, b, w, k);
    }
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ltrsv, dim3(1), dim3(1024), static_cast<size_t>(n) * sizeof(double), s, b, w);
  hipLaunchKernelGGL(k_lepilogue, dim3(1), dim3(256), 0, s, b, w);
  return hipGetLastError();
}

}  // namespace sba_dev
