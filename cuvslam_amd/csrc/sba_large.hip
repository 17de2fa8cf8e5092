// Reduced camera system of windows with more than kMaxFreePoses free key frames ("mapped" windows, e.g. a large
// multi-camera map with a reduced system of order ~1200).
//
// The Schur kernel stages at most kMaxFreePoses free key frames per chunk; a mapped window's chunks carry the
// list of the free key frames they stage (chunk_map, increasing). Here:
//   k_gather      A = U + lambda diag(U) - sum Z Z^T in FP64 (dense, row-major lower triangle), rhs e: each
//                 key-frame block sums its contributions from the chunk records in chunk order (a plan built
//                 with the chunks), so the system, like the small windows' reduction, has fixed bits
//   k_potrf / k_trsm / k_syrk   right-looking blocked Cholesky (Golub & Van Loan, Matrix Computations,
//                 Algorithm 4.2.10) with 64 x 64 blocks: the diagonal block is factored with the 16 x 16 FP64
//                 tile scheme of the small-window factor (matrix cores for the tile updates) and its triangular
//                 inverse kept; the panel below it is then a product with that inverse (one block per
//                 workgroup), the trailing update one block pair per workgroup.
//                 The forward substitution L y = e rides along: y_k = L_kk^-1 e_k in k_potrf, e_i -= L_ik y_k
//                 in k_trsm
//   k_bsolve      L^T x = y by 64-row blocks with the kept inverses, in one workgroup
//   k_lepilogue   the pose step, its max-norm and dc^T diag(U) dc, trial key frames
// A non-positive pivot clears the window's chol_ok; the later kernels then return, and the LM decision raises
// lambda (schur_complement_bundler_gpu.cpp:175-185), as for the small windows.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "sba_device.h"
#include "sba_tiles.h"

using namespace sba_dev;
using namespace sba_tiles;

namespace {

constexpr int kLB = 64;        // block of the large factorization
constexpr int kGather = 8;     // contributions whose loads k_gather issues together
constexpr int kGatherWaves = 4;  // wavefronts of k_gather that share one key-frame block's contributions

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

}  // namespace

// A = U + lambda diag(U) - sum Z Z^T (lower triangle), rhs = e, du = diag(U); chol_ok set for the factor.
// One workgroup per key-frame block (a, b), a <= b (blockIdx.x enumerates the upper blocks row by row). Lane 6 r + c
// of each of the kGatherWaves wavefronts sums element (r, c) of the chunk-local block (la, lb) (36 contiguous
// floats of a record) in FP64 over its share of the block's contributions, in their listed (chunk) order; the
// shares' sums are added in wavefront order.
// A(6 b + c, 6 a + r) = -sum; in a diagonal block lanes with r <= c add U and lambda diag U, lanes 36..41 sum e
// and lanes 42..47 diag U. The sums do not depend on the order the hardware runs the chunks or the workgroups.
__global__ __launch_bounds__(64 * kGatherWaves) void k_gather(Batch b, int w) {
  __shared__ double red[3][kGatherWaves - 1][64];  // wavefronts 1..: their sums z, u, d
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done) return;
  const int ncf = W.ncf, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // upper block t = blockIdx.x: row ka holds blocks ka..ncf-1, rows before it ka ncf - ka (ka - 1) / 2 blocks
  const int t = blockIdx.x;
  int ka = static_cast<int>((2.0 * ncf + 1.0 - sqrt((2.0 * ncf + 1.0) * (2.0 * ncf + 1.0) - 8.0 * t)) / 2.0);
  ka = max(0, min(ka, ncf - 1));
  while (ka > 0 && ka * ncf - (ka * (ka - 1)) / 2 > t) --ka;
  while (ka + 1 < ncf && (ka + 1) * ncf - ((ka + 1) * ka) / 2 <= t) ++ka;
  const int kb = ka + t - (ka * ncf - (ka * (ka - 1)) / 2);
  if (t == 0 && threadIdx.x == 0) St->chol_ok = 1;
  const Large L = large_of(b, W);
  const int n = L.n;
  const bool diag = ka == kb;
  const int r = lane / 6, c = lane % 6;
  // what this lane sums: 0 Z Z^T element (with U and diag U on the diagonal), 1 e, 2 diag U, -1 nothing
  int kind = -1;
  if (lane < 36 && (!diag || r <= c)) kind = 0;
  else if (diag && lane >= 36 && lane < 48) kind = lane < 42 ? 1 : 2;
  const int q = lane >= 36 ? (lane - 36) % 6 : 0;
  const int32_t* first = b.gfirst + W.gfirst_off + ka * ncf + kb;
  const int nall = first[1] - first[0];
  const int kbeg = first[0] + (nall * wave) / kGatherWaves, kend = first[0] + (nall * (wave + 1)) / kGatherWaves;
  const bool want_u = kind == 0 && diag, want_d = want_u && r == c;
  double z = 0.0, u = 0.0, d = 0.0;
  // kGather contributions at a time: their plan entries, then their record values, all loads of a batch in
  // flight together; the values are then added one after another in the plan's order
  for (int k0 = kbeg; kind >= 0 && k0 < kend; k0 += kGather) {
    int64_t oz[kGather], ou[kGather], od[kGather];
#pragma unroll
    for (int t = 0; t < kGather; ++t) {
      const int4 g = b.gcon[min(k0 + t, kend - 1)];
      const int64_t rec = (static_cast<int64_t>(g.w) << 32) | static_cast<uint32_t>(g.z);
      const int la = g.y & 255, lb = (g.y >> 8) & 255, nloc = g.y >> 24;
      const int64_t tail = rec + 36 * (nloc * (nloc + 1) / 2);
      if (kind == 0) {
        oz[t] = rec + 36 * (la * nloc - (la * (la - 1)) / 2 + (lb - la)) + 6 * r + c;
      } else {
        oz[t] = tail + (kind == 1 ? 21 : 27) * nloc + 6 * la + q;
      }
      ou[t] = tail + 21 * la + (want_u ? 6 * r - (r * (r - 1)) / 2 + (c - r) : 0);
      od[t] = tail + 27 * nloc + 6 * la + r;
    }
    float vz[kGather], vu[kGather], vd[kGather];
#pragma unroll
    for (int t = 0; t < kGather; ++t) {
      vz[t] = b.part[oz[t]];
      vu[t] = want_u ? b.part[ou[t]] : 0.f;
      vd[t] = want_d ? b.part[od[t]] : 0.f;
    }
#pragma unroll
    for (int t = 0; t < kGather; ++t)
      if (k0 + t < kend) {
        z += static_cast<double>(vz[t]);
        u += static_cast<double>(vu[t]);
        d += static_cast<double>(vd[t]);
      }
  }
  if (wave > 0) {
    red[0][wave - 1][lane] = z;
    red[1][wave - 1][lane] = u;
    red[2][wave - 1][lane] = d;
  }
  __syncthreads();
  if (wave > 0 || kind < 0) return;
#pragma unroll
  for (int v = 0; v < kGatherWaves - 1; ++v) {
    z += red[0][v][lane];
    u += red[1][v][lane];
    d += red[2][v][lane];
  }
  if (kind == 0) {
    double v = u - z;
    if (diag && r == c) v += static_cast<double>(St->lambda) * d;
    L.A[static_cast<int64_t>(6 * kb + c) * n + 6 * ka + r] = v;
  } else if (kind == 1) {
    L.rhs[6 * ka + q] = z;
  } else {
    L.du[6 * ka + q] = z;
  }
}

// Cholesky of diagonal block k (64 x 64; rows and columns past the matrix are identity) by one workgroup of 256
// threads, as four 16 x 16 tile columns: wavefront 0 factors the diagonal tiles (kept in LDS) one after another
// (factor_tile), wavefronts 1..3 hold the six off-diagonal tiles U_JI (J < I, A = U^T U) in registers and apply
// the panel step U_KI = X_KK A_KI and the trailing updates on the FP64 matrix cores; the forward-substitution
// step y_k = L_kk^-1 e_k runs alongside. Then the block's inverse X = L_kk^-1 (for the panel GEMMs and the
// backward solve), tile by tile: X_IJ = -X_II sum_{K=J}^{I-1} L_IK X_KJ, one tile column J per wavefront 1..3.
__global__ __launch_bounds__(256) void k_potrf(Batch b, int w, int k) {
  constexpr int T = kLB / 16;       // 4 tile columns
  __shared__ double sDg[T * 256];   // diagonal tiles, row-major (U after factoring)
  __shared__ double sXT[T * 256];   // X_KK^T per tile column (row-major)
  __shared__ double sP[T * 256];    // panel tiles U_KI of the current tile column
  __shared__ double sU[6 * 256];    // off-diagonal U tiles (row-major), by pair index
  __shared__ double sX[6 * 256];    // off-diagonal X tiles X_IJ (row-major), by the pair index of (J, I)
  __shared__ double sS[3 * 256];    // per-wavefront product scratch
  __shared__ double sT[256];
  __shared__ double vec[kLB], ys[kLB], idg[kLB];
  __shared__ int flag;
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  const int lane = tid & 63, li = lane >> 4, lj = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // pair index of tile (J, I), J < I: 0 (0,1) 1 (0,2) 2 (0,3) 3 (1,2) 4 (1,3) 5 (2,3)
  auto pair_of = [](int J, int I) { return J == 0 ? I - 1 : (J == 1 ? I + 1 : 5); };
  // element (4v + lane/16, lane%16) of upper tile (J, I) of the symmetric block, read from the lower triangle
  auto element = [&](int J, int I, int v) -> double {
    const int r = 16 * J + 4 * v + li, c = 16 * I + lj;
    if (r >= kb || c >= kb) return r == c ? 1.0 : 0.0;
    return L.A[static_cast<int64_t>(k0 + max(r, c)) * n + k0 + min(r, c)];
  };
  {
    double val[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) val[v] = element(wave, wave, v);  // diagonal tile `wave`
#pragma unroll
    for (int v = 0; v < 4; ++v) sDg[256 * wave + 64 * v + lane] = val[v];
  }
  if (tid < kLB) vec[tid] = tid < kb ? L.rhs[k0 + tid] : 0.0;
  if (tid == 0) flag = 0;

  if (wave == 0) {
    __syncthreads();
    for (int K = -1; K < T; ++K) {
      if (flag) break;
      if (K >= 0) {
        const double* xt = sXT + 256 * K;
        double yk = 0.0;  // y_K = X_KK e_K
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) yk += xt[16 * (4 * li + kk) + lj] * vec[16 * K + 4 * li + kk];
        yk = sum_rows4(yk);
        if (lane < 16) ys[16 * K + lane] = yk;
        __syncthreads();
      }
      if (K + 1 < T) {
        double* tk = sDg + 256 * (K + 1);
        double4v dt;
#pragma unroll
        for (int v = 0; v < 4; ++v) dt[v] = tk[64 * v + lane];
        if (K >= 0) dt = tile_update(dt, sP + 256 * (K + 1), sP + 256 * (K + 1), lane);
        if (!factor_tile(dt, sT, sXT + 256 * (K + 1), idg + 16 * (K + 1), lane) && lane == 0) flag = 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) tk[64 * v + lane] = dt[v];
      }
      __syncthreads();
    }
    __syncthreads();  // the off-diagonal U tiles are published
    __syncthreads();  // the off-diagonal X tiles are formed
  } else {
    // two register tiles per wavefront: pair index m = 3 s + wave - 1
    int tJ[2], tI[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int m = 3 * s + wave - 1;
      tJ[s] = m < 3 ? 0 : (m < 5 ? 1 : 2);
      tI[s] = m < 3 ? m + 1 : (m < 5 ? m - 1 : 3);
    }
    double4v acc[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[s][v] = element(tJ[s], tI[s], v);
    __syncthreads();
    for (int K = -1; K < T; ++K) {
      if (flag) break;
      if (K >= 0) {
        // panel U_KI = X_KK A_KI (the B operand of chunk q is register q of the tile), e_I -= U_KI^T y_K
        const double* xt = sXT + 256 * K;
        double yk = 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) yk += xt[16 * (4 * li + kk) + lj] * vec[16 * K + 4 * li + kk];
        yk = sum_rows4(yk);
        double yv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) yv[v] = __shfl(yk, 4 * v + li);  // y_K[4v + lane/16]
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (tJ[s] == K) {
            const int I = tI[s];
            double4v u = double4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) u = mfma64(xt[64 * q + lane], acc[s][q], u);
            acc[s] = u;
            double p = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              sP[256 * I + 64 * v + lane] = u[v];
              p += u[v] * yv[v];
            }
            p = sum_rows4(p);
            if (lane < 16) vec[16 * I + lane] -= p;
          }
        }
        __syncthreads();
        // trailing: the diagonal tiles past K + 1 (LDS), then this wavefront's tiles (J > K)
        for (int J = K + 1 + wave; J < T; J += 3) {
          double* tj = sDg + 256 * J;
          double4v t;
#pragma unroll
          for (int v = 0; v < 4; ++v) t[v] = tj[64 * v + lane];
          t = tile_update(t, sP + 256 * J, sP + 256 * J, lane);
#pragma unroll
          for (int v = 0; v < 4; ++v) tj[64 * v + lane] = t[v];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
          if (tJ[s] > K) acc[s] = tile_update(acc[s], sP + 256 * tJ[s], sP + 256 * tI[s], lane);
      }
      __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v) sU[256 * pair_of(tJ[s], tI[s]) + 64 * v + lane] = acc[s][v];
    __syncthreads();
    if (flag == 0) {
      // tile column J = wave - 1 of X: X_IJ = -X_II S, S = sum_{K=J}^{I-1} L_IK X_KJ, L_IK = U_KI^T
      const int J = wave - 1;
      double* S = sS + 256 * J;
      double4v xo[3];
#pragma unroll
      for (int I = 1; I < T; ++I) {
        if (I <= J) continue;
        double4v sa = double4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int K = 0; K < T - 1; ++K) {
          if (K < J || K >= I) continue;
          const double* U = sU + 256 * pair_of(K, I);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            // A operand L_IK[l%16][4q + l/16] = U_KI[4q + l/16][l%16]; B operand X_KJ[4q + l/16][l%16]: from
            // X_JJ^T in LDS, or register q of the tile X_KJ formed above
            const double xb = K == J ? sXT[256 * J + 16 * lj + 4 * q + li] : xo[K - J - 1][q];
            sa = mfma64(U[64 * q + lane], xb, sa);
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) S[64 * v + lane] = sa[v];
        wave_lds_sync();
        // A operand X_II[l%16][4q + l/16] = X_II^T[4q + l/16][l%16]
        double4v d = double4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q) d = mfma64(-sXT[256 * I + 64 * q + lane], S[64 * q + lane], d);
        xo[I - J - 1] = d;
#pragma unroll
        for (int v = 0; v < 4; ++v) sX[256 * pair_of(J, I) + 64 * v + lane] = d[v];
        wave_lds_sync();  // S is rewritten for the next I
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (flag != 0) {
    if (tid == 0) St->chol_ok = 0;
    return;
  }
  // L (lower) back to A, X = L_kk^-1 (64 x 64 row-major, zero above the diagonal) to Linv, y_k to rhs.
  // U tile (J, I) element (r, c) is L[16 I + c][16 J + r]; X_II = (X_II^T)^T from sXT.
  double* Xg = L.Linv + static_cast<int64_t>(k) * kLB * kLB;
  for (int idx = tid; idx < kLB * kLB; idx += blockDim.x) {
    const int R = idx / kLB, Cc = idx % kLB;
    const int I = R >> 4, J = Cc >> 4, a = R & 15, c = Cc & 15;
    double x = 0.0;
    if (J == I) x = sXT[256 * I + 16 * c + a];
    else if (J < I) x = sX[256 * pair_of(J, I) + 16 * a + c];
    Xg[idx] = x;
    if (Cc <= R && R < kb) {
      const double l = J == I ? sDg[256 * I + 16 * c + a] : sU[256 * pair_of(J, I) + 16 * c + a];
      L.A[static_cast<int64_t>(k0 + R) * n + k0 + Cc] = l;
    }
  }
  if (tid < kb) L.rhs[k0 + tid] = ys[tid];
}

// 64 x 64 products C -= P Q^T on the FP64 matrix cores for the panel and trailing kernels below: P and Q are staged
// in LDS as transposed 16 x 16 tiles (tile (I, M) of P at sPT + 256 (4 I + M), element [m][i] = P[16 I + i][16 M + m],
// the operand layout of tile_update), and each of the four wavefronts owns four of the 16 output tiles (I, J)
// = (wave, 0..3), accumulated over the four tile columns M in register tiles.
// (256 threads: each loads its 16 entries in one batch, then stores them.)
__device__ __forceinline__ void stage_transposed(double* sT, const double* src, int64_t ld, int rows, int cols,
                                                 int tid) {
  constexpr int kPer = kLB * kLB / 256;
  double v[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int idx = tid + 256 * it, r = idx / kLB, c = idx % kLB;  // row-major: consecutive threads, columns
    v[it] = (r < rows && c < cols) ? src[static_cast<int64_t>(r) * ld + c] : 0.0;
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int idx = tid + 256 * it, r = idx / kLB, c = idx % kLB;
    sT[256 * (4 * (r >> 4) + (c >> 4)) + 16 * (c & 15) + (r & 15)] = v[it];
  }
}

__device__ __forceinline__ void gemm_nt_update(double4v (&acc)[4], const double* sPT, const double* sQT, int wave,
                                               int lane) {
#pragma unroll
  for (int J = 0; J < 4; ++J)
#pragma unroll
    for (int M = 0; M < 4; ++M)
      acc[J] = tile_update(acc[J], sPT + 256 * (4 * wave + M), sQT + 256 * (4 * J + M), lane);
}

// Panel below diagonal block k: L_ik = A_ik L_kk^-T = A_ik X^T with the kept inverse X, one block row per
// workgroup, on the matrix cores; then the block row's forward-substitution update e_i -= L_ik y_k.
__global__ __launch_bounds__(256) void k_trsm(Batch b, int w, int k) {
  __shared__ double sAT[16 * 256];
  __shared__ double sXT[16 * 256];
  __shared__ double sy[kLB];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i0 = kLB * (k + 1 + blockIdx.x), ib = min(kLB, n - i0);
  const double* X = L.Linv + static_cast<int64_t>(k) * kLB * kLB;
  stage_transposed(sAT, L.A + static_cast<int64_t>(i0) * n + k0, n, ib, kb, tid);
  stage_transposed(sXT, X, kLB, kLB, kLB, tid);
  if (tid < kLB) sy[tid] = tid < kb ? L.rhs[k0 + tid] : 0.0;
  __syncthreads();
  double4v acc[4];
#pragma unroll
  for (int J = 0; J < 4; ++J) acc[J] = double4v{0.0, 0.0, 0.0, 0.0};
  gemm_nt_update(acc, sAT, sXT, wave, lane);  // -(A X^T) for the tiles (wave, J)
  __syncthreads();  // every wavefront has read sAT before it is overwritten with L_ik (row-major)
  double* sL = sAT;
#pragma unroll
  for (int J = 0; J < 4; ++J)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * wave + 4 * v + (lane >> 4), c = 16 * J + (lane & 15);
      const double val = (r < ib && c < kb) ? -acc[J][v] : 0.0;
      sL[r * kLB + c] = val;
      if (r < ib && c < kb) L.A[static_cast<int64_t>(i0 + r) * n + k0 + c] = val;
    }
  __syncthreads();
  if (tid < ib) {
    double s = 0.0;
    for (int m = 0; m < kb; ++m) s += sL[tid * kLB + m] * sy[m];
    L.rhs[i0 + tid] -= s;
  }
}

// Trailing update after block k: A_ij -= L_ik L_jk^T for k < j <= i, one block pair per workgroup, on the matrix
// cores (the lower triangle of a diagonal block).
__global__ __launch_bounds__(256) void k_syrk(Batch b, int w, int k) {
  __shared__ double sAT[16 * 256];
  __shared__ double sBT[16 * 256];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, k0 = kLB * k, kb = min(kLB, n - k0), tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  int ii = 0, t = blockIdx.x;
  while (t > ii) {
    t -= ii + 1;
    ++ii;
  }
  const int jj = t;  // 0 <= jj <= ii
  const int i0 = kLB * (k + 1 + ii), j0 = kLB * (k + 1 + jj);
  const int ib = min(kLB, n - i0), jb = min(kLB, n - j0);
  stage_transposed(sAT, L.A + static_cast<int64_t>(i0) * n + k0, n, ib, kb, tid);
  stage_transposed(sBT, L.A + static_cast<int64_t>(j0) * n + k0, n, jb, kb, tid);
  // this wavefront's output tiles (wave, J), element (4 v + lane / 16, lane % 16) in register v
  double4v acc[4];
#pragma unroll
  for (int J = 0; J < 4; ++J)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * wave + 4 * v + (lane >> 4), c = 16 * J + (lane & 15);
      acc[J][v] = (r < ib && c < jb) ? L.A[static_cast<int64_t>(i0 + r) * n + j0 + c] : 0.0;
    }
  __syncthreads();
  gemm_nt_update(acc, sAT, sBT, wave, lane);
#pragma unroll
  for (int J = 0; J < 4; ++J)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * wave + 4 * v + (lane >> 4), c = 16 * J + (lane & 15);
      if (r < ib && c < jb && (ii != jj || c <= r)) L.A[static_cast<int64_t>(i0 + r) * n + j0 + c] = acc[J][v];
    }
}

// Backward substitution L^T x = y in one workgroup of 1024 threads, 64-row blocks from the last: for block bk,
// r = y_bk - sum_{i past the block} L[i][block bk]^T x_i (16 groups of 64 threads, each group one row at a time
// with its 64 lanes on the block's 64 contiguous columns), then x_bk = X_bk^T r with the kept inverse. x stays
// in the dynamic shared memory (n doubles) and goes to rhs at the end.
__global__ __launch_bounds__(1024) void k_bsolve(Batch b, int w) {
  constexpr int kG = 16;                        // row groups
  extern __shared__ double sx[];                 // y, overwritten by x block by block
  __shared__ double part[kG * kLB];
  __shared__ double sr[kLB];
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || !St->chol_ok) return;
  const Large L = large_of(b, W);
  const int n = L.n, tid = threadIdx.x, c = tid & 63, g = tid >> 6;
  const int nb = (n + kLB - 1) / kLB;
  for (int i = tid; i < n; i += blockDim.x) sx[i] = L.rhs[i];
  __syncthreads();
  for (int bk = nb - 1; bk >= 0; --bk) {
    const int b0 = kLB * bk, bb = min(kLB, n - b0);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;  // four rows in flight per thread
    if (c < bb) {
      const double* col = L.A + b0 + c;
      int i = b0 + kLB + g;
      for (; i + 3 * kG < n; i += 4 * kG) {
        s0 += col[static_cast<int64_t>(i) * n] * sx[i];
        s1 += col[static_cast<int64_t>(i + kG) * n] * sx[i + kG];
        s2 += col[static_cast<int64_t>(i + 2 * kG) * n] * sx[i + 2 * kG];
        s3 += col[static_cast<int64_t>(i + 3 * kG) * n] * sx[i + 3 * kG];
      }
      for (; i < n; i += kG) s0 += col[static_cast<int64_t>(i) * n] * sx[i];
    }
    part[g * kLB + c] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (tid < kLB) {
      double r = 0.0;
      if (c < bb) {
        r = sx[b0 + c];
#pragma unroll
        for (int q = 0; q < kG; ++q) r -= part[q * kLB + c];
      }
      sr[c] = r;
    }
    __syncthreads();
    // x_c = sum_{m >= c} X[m][c] r_m, the sum split over the 16 groups
    const double* X = L.Linv + static_cast<int64_t>(bk) * kLB * kLB;
    double xs = 0.0;
    if (c < bb)
      for (int m = c + g; m < bb; m += kG) xs += X[m * kLB + c] * sr[m];
    part[g * kLB + c] = xs;
    __syncthreads();
    if (tid < bb) {
      double x = 0.0;
#pragma unroll
      for (int q = 0; q < kG; ++q) x += part[q * kLB + c];
      sx[b0 + c] = x;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) L.rhs[i] = sx[i];
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

// The factor and solve of window w (n = 6 free key frames): gather, blocked Cholesky, triangular solves,
// epilogue. 3 nb launches for nb 64-row blocks.
hipError_t launch_large_factor(const Batch& b, int w, int n, hipStream_t s) {
  const int nb = (n + kLB - 1) / kLB, ncf = n / 6;
  hipLaunchKernelGGL(k_gather, dim3(ncf * (ncf + 1) / 2), dim3(64 * kGatherWaves), 0, s, b, w);
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
  hipLaunchKernelGGL(k_bsolve, dim3(1), dim3(1024), sizeof(double) * n, s, b, w);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lepilogue, dim3(1), dim3(256), 0, s, b, w);
  return hipGetLastError();
}

}  // namespace sba_dev
