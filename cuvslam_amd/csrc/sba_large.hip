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
//                 Algorithm 4.2.10) with 64 x 64 blocks: the diagonal block is factored in one wavefront with
//                 its rows in registers, and its triangular inverse kept; the panel below it is then a product
//                 with that inverse (one block per workgroup), the trailing update one block pair per workgroup.
//                 The forward substitution L y = e rides along: y_k = L_kk^-1 e_k in k_potrf, e_i -= L_ik y_k
//                 in k_trsm
//   k_bsolve      L^T x = y, one launch per 64-row block with the kept inverses
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

struct Large {
  double* A;     // n x n, row-major; the lower triangle holds A, then L
  double* rhs;   // e, then y, then x
  double* du;    // diag(U)
  double* Linv;  // per 64-row block: the inverse of its diagonal block of L (64 x 64, row-major, lower)
  int n;
};

__device__ __forceinline__ Large large_of(const Batch& b, const WindowDesc& W) {
  const int n = 6 * W.ncf;
  double* A = b.A + W.a_off;
  const int64_t nn = static_cast<int64_t>(n) * n;
  return Large{A, A + nn, A + nn + n, A + nn + 2 * n, n};
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

// Cholesky of diagonal block k by one workgroup of 256 threads, with its inverse X = L_kk^-1 and the forward-
// substitution step y_k = X e_k formed alongside (Golub & Van Loan 4.2 and 3.1, column-oriented). Thread
// (wavefront v, lane c) keeps rows r = 4t + v (t < 16) of column c of both the block and X in registers. Per panel
// of 4 columns: the panel columns and the panel rows of X go to LDS; wavefront 0 factors the panel (lane = row,
// pivots and column entries broadcast with readlane) and applies it to the X rows and to e (lane = column); then
// every thread applies the rank-4 update to its cells of the trailing block, of X and of e. LDS buffers
// alternate between panels, so two barriers per panel suffice. Rows and columns past the matrix are identity.
__global__ __launch_bounds__(256) void k_potrf(Batch b, int w, int k) {
  constexpr int kP = 4, kT = kLB / 4;  // panel width (= wavefronts per workgroup), rows per thread
  __shared__ double sP[2][kLB][kP];    // panel columns, [row][q]
  __shared__ double sX[2][kP][kLB];    // panel rows of X, [q][column]
  __shared__ double sE[kLB];           // e_k, overwritten with y_k
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  double a[kT], x[kT];
#pragma unroll
  for (int t = 0; t < kT; ++t) {
    const int r = 4 * t + wv;
    double v = 0.0;
    if (r < kb && lane < kb) {
      if (lane <= r) v = L.A[static_cast<int64_t>(k0 + r) * n + k0 + lane];
    } else if (r == lane) {
      v = 1.0;
    }
    a[t] = v;
    x[t] = r == lane ? 1.0 : 0.0;
  }
  if (tid < kLB) sE[tid] = tid < kb ? L.rhs[k0 + tid] : 0.0;
  bool ok = true;
  for (int p = 0; p < kLB / kP; ++p) {
    const int c0 = kP * p, buf = p & 1;
    if (lane >= c0 && lane < c0 + kP) {  // rows above the diagonal hold zeros
#pragma unroll
      for (int t = 0; t < kT; ++t) sP[buf][4 * t + wv][lane - c0] = a[t];
    }
    {
      double v = x[0];  // X row c0 + wv is x[p] of wavefront wv
#pragma unroll
      for (int t = 1; t < kT; ++t) v = t == p ? x[t] : v;
      sX[buf][wv][lane] = v;
    }
    __syncthreads();
    if (wv == 0) {
      double c[kP], rl[kP], lpp[kP][kP];
#pragma unroll
      for (int q = 0; q < kP; ++q) c[q] = lane >= c0 ? sP[buf][lane][q] : 0.0;
#pragma unroll
      for (int q = 0; q < kP; ++q) {
        const int cq = c0 + q;
        const double d = readlane_d(c[q], cq);
        if (!(d > 0.0)) ok = false;
        const double l = sqrt(fmax(d, 1e-300));
        rl[q] = 1.0 / l;
        c[q] = lane == cq ? l : (lane > cq ? c[q] * rl[q] : 0.0);
#pragma unroll
        for (int q2 = q + 1; q2 < kP; ++q2) {
          lpp[q2][q] = readlane_d(c[q], c0 + q2);  // L[c0 + q2][c0 + q]
          if (lane >= c0 + q2) c[q2] -= c[q] * lpp[q2][q];
        }
      }
#pragma unroll
      for (int q = 0; q < kP; ++q) sP[buf][lane][q] = c[q];  // zero above the diagonal
      double xr[kP], y[kP];
#pragma unroll
      for (int q = 0; q < kP; ++q) {
        double v = sX[buf][q][lane], u = sE[c0 + q];
#pragma unroll
        for (int q2 = 0; q2 < q; ++q2) {
          v -= lpp[q][q2] * xr[q2];
          u -= lpp[q][q2] * y[q2];
        }
        xr[q] = v * rl[q];
        y[q] = u * rl[q];
        sX[buf][q][lane] = xr[q];
      }
      // e_r -= L[r][panel] y[panel] for the rows below the panel (lane = row)
      double de = 0.0;
#pragma unroll
      for (int q = 0; q < kP; ++q) de += c[q] * y[q];
      if (lane >= c0 + kP) sE[lane] -= de;
      else if (lane >= c0) sE[lane] = lane == c0 ? y[0] : lane == c0 + 1 ? y[1] : lane == c0 + 2 ? y[2] : y[3];
    }
    __syncthreads();
    // all loads first: this thread's column of the panel, X's panel rows, and the panel entries of its rows
    double lc[kP], xr[kP], lr[kT][kP];
#pragma unroll
    for (int q = 0; q < kP; ++q) {
      lc[q] = sP[buf][lane][q];
      xr[q] = sX[buf][q][lane];
    }
#pragma unroll
    for (int t = 0; t < kT; ++t)
#pragma unroll
      for (int q = 0; q < kP; ++q) lr[t][q] = sP[buf][4 * t + wv][q];
    // Branch-free over the rows: panel entries of rows above the panel are zero, so those rows do not change.
    const bool below = lane >= c0 + kP, in_panel = lane >= c0 && !below;
    const int qp = lane - c0;
    const double xown = wv == 0 ? xr[0] : wv == 1 ? xr[1] : wv == 2 ? xr[2] : xr[3];
#pragma unroll
    for (int t = 0; t < kT; ++t) {
      const int r = 4 * t + wv;
      double da = 0.0, dx = 0.0;
#pragma unroll
      for (int q = 0; q < kP; ++q) {
        da += lr[t][q] * lc[q];
        dx += lr[t][q] * xr[q];
      }
      const double lsel = sP[buf][r][qp & (kP - 1)];  // L[r][lane] for the panel's lanes
      if (lane <= r) a[t] = below ? a[t] - da : (in_panel ? lsel : a[t]);
      x[t] = (r >= c0 && r < c0 + kP) ? xown : x[t] - dx;  // row c0 + wv of X is final
    }
  }
#pragma unroll
  for (int t = 0; t < kT; ++t) {
    const int r = 4 * t + wv;
    if (r < kb && lane <= r) L.A[static_cast<int64_t>(k0 + r) * n + k0 + lane] = a[t];
    L.Linv[static_cast<int64_t>(k) * kLB * kLB + r * kLB + lane] = x[t];
  }
  __syncthreads();
  if (tid < kb) L.rhs[k0 + tid] = sE[tid];
  if (tid == 0 && !ok) St->chol_ok = 0;
}

// Panel below diagonal block k: L_ik = A_ik L_kk^-T = A_ik X^T with the kept inverse X, one block row per
// workgroup (each thread a 4 x 4 piece); then the block row's forward-substitution update e_i -= L_ik y_k.
__global__ __launch_bounds__(256) void k_trsm(Batch b, int w, int k) {
  __shared__ double sA[kLB * kLds];
  __shared__ double sX[kLB * kLds];
  __shared__ double sy[kLB];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  const int i0 = kLB * (k + 1 + blockIdx.x), ib = min(kLB, n - i0);
  const double* X = L.Linv + static_cast<int64_t>(k) * kLB * kLB;
  for (int idx = tid; idx < kLB * kLB; idx += blockDim.x) {
    const int r = idx / kLB, c = idx % kLB;
    sA[r * kLds + c] = (r < ib && c < kb) ? L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] : 0.0;
    sX[r * kLds + c] = X[idx];
  }
  if (tid < kLB) sy[tid] = tid < kb ? L.rhs[k0 + tid] : 0.0;
  __syncthreads();
  const int r0 = 4 * (tid >> 4), c0 = 4 * (tid & 15);
  double acc[4][4] = {};
  for (int m = 0; m < kLB; ++m) {
    double a[4], x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = sA[(r0 + q) * kLds + m];
      x[q] = sX[(c0 + q) * kLds + m];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * x[q];
  }
  __syncthreads();  // sA is read by every thread before it is overwritten with L_ik
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + p, c = c0 + q;
      sA[r * kLds + c] = (r < ib && c < kb) ? acc[p][q] : 0.0;
      if (r < ib && c < kb) L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] = acc[p][q];
    }
  __syncthreads();
  if (tid < ib) {
    double s = 0.0;
    for (int m = 0; m < kb; ++m) s += sA[tid * kLds + m] * sy[m];
    L.rhs[i0 + tid] -= s;
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

// Backward substitution L^T x = y, one launch per 64-row block from the last (y from the factor kernels; rhs
// holds w = y minus the products with the solved blocks). Launch for block bk, one workgroup per block j <= bk:
// each computes x_bk = X_bk^T w_bk from the kept inverse (64 x 64, cheap to repeat); workgroup bk stores it and
// every other workgroup j < bk subtracts L_bk,j^T x_bk from w_j. Block bk is final once its launch has run.
__global__ __launch_bounds__(64) void k_bsolve(Batch b, int w, int bk) {
  __shared__ double sw[kLB];
  __shared__ double sx[kLB];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, lane = threadIdx.x, j = blockIdx.x;
  const int b0 = kLB * bk, bb = min(kLB, n - b0);
  sw[lane] = lane < bb ? L.rhs[b0 + lane] : 0.0;
  __syncthreads();
  const double* X = L.Linv + static_cast<int64_t>(bk) * kLB * kLB;
  double x = 0.0;
  for (int m = lane; m < bb; ++m) x += X[m * kLB + lane] * sw[m];  // x_c = sum_{m >= c} X[m][c] w_m
  if (j == bk) {
    if (lane < bb) L.rhs[b0 + lane] = x;
    return;
  }
  sx[lane] = lane < bb ? x : 0.0;
  __syncthreads();
  // w_j[c] -= sum_r L[b0 + r][j0 + c] x_r (rows of block bk, columns of block j < bk)
  const int j0 = kLB * j;
  double s = 0.0;
  for (int r = 0; r < bb; ++r) s += L.A[static_cast<int64_t>(b0 + r) * n + j0 + lane] * sx[r];
  L.rhs[j0 + lane] -= s;
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
  for (int bk = nb - 1; bk >= 0 && e == hipSuccess; --bk) {
    hipLaunchKernelGGL(k_bsolve, dim3(bk + 1), dim3(64), 0, s, b, w, bk);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lepilogue, dim3(1), dim3(256), 0, s, b, w);
  return hipGetLastError();
}

}  // namespace sba_dev
