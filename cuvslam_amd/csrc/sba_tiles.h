// Device helpers shared by the FP64 factor kernels: 16 x 16 FP64 tiles in the matrix-core layout
// (v_mfma_f64_16x16x4_f64: element (4v + lane/16, lane%16) in register v of the lane, i.e. 64 v + lane in
// row-major order), their trailing update, the one-wavefront Cholesky of a diagonal tile, and lane reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace sba_tiles {

// Makes the LDS writes of some lanes of a wavefront visible to the later LDS reads of the other lanes: without
// it the compiler may satisfy a lane's read before the stores that only other lanes execute.
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

typedef double double4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double readlane_d(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}


__device__ __forceinline__ double4v mfma64(double a, double b, double4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// c - P_J^T P_I for two 16 x 16 row-major tiles in LDS (A operand lane l: P_J[4q + l/16][l%16] = 64 q + l)
__device__ __forceinline__ double4v tile_update(double4v c, const double* pj, const double* pi, int lane) {
#pragma unroll
  for (int q = 0; q < 4; ++q) c = mfma64(-pj[64 * q + lane], pi[64 * q + lane], c);
  return c;
}

// A wavefront-uniform value the compiler may not hoist out of the enclosing loop: keeps the per-tile LDS
// addresses from being precomputed for every tile slot at once (which would cost a register per slot).
__device__ __forceinline__ int in_loop(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

__device__ __forceinline__ double sum_rows4(double x) {  // over lanes l, l^16, l^32, l^48
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

__device__ __forceinline__ double sum_row16(double x) {  // over the 16 lanes of l's row group
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Cholesky of a symmetric 16 x 16 tile held by one wavefront in the FP64 matrix-core layout (cell
// (4v + lane/16, lane%16) in t[v]). The tile goes through LDS (sT, row-major) to rows in lanes: lane i < 16 keeps
// row i of L, pivots and column entries broadcast with readlane. Then X = L^-1 by forward substitution, lane j
// solving column j against L read back from LDS. On return t holds U = L^T (zero below the diagonal), X^T is
// written to xt (row-major) and 1 / L_ii to rdiag. False on a non-positive pivot.
__device__ inline bool factor_tile(double4v& t, double* sT, double* xt, double* rdiag, int lane) {
  const int i = lane & 15;
#pragma unroll
  for (int v = 0; v < 4; ++v) sT[64 * v + lane] = t[v];
  wave_lds_sync();
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = sT[16 * i + q];
  wave_lds_sync();  // every row read before sT is rewritten
  bool ok = true;
  double ri = 0.0;  // 1 / L_ii of this lane's row
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const double d = readlane_d(a[c], c);
    if (!(d > 0.0)) ok = false;
    // 1/sqrt(d) by the hardware estimate refined with one Newton-Raphson step, L_cc = d / sqrt(d)
    const double dd = fmax(d, 1e-300);
    double il = __builtin_amdgcn_rsq(dd);
    il = il * (1.5 - 0.5 * dd * il * il);
    a[c] = i == c ? dd * il : (i > c ? a[c] * il : 0.0);
    ri = i == c ? il : ri;
    // rows i < c2 update their a[c2] too: it lies above the diagonal there, and pivot c2 sets it to zero. Column
    // c + 1 first: the next pivot depends only on it, so its broadcast and square root can start while the
    // other columns are updated.
#pragma unroll
    for (int c2 = c + 1; c2 < 16; ++c2) a[c2] -= a[c] * readlane_d(a[c], c2);  // L[c2][c]
  }
  if (lane < 16) {
    rdiag[lane] = ri;
#pragma unroll
    for (int q = 0; q < 16; ++q) sT[16 * q + i] = a[q];  // U[q][i] = L[i][q], zero for q > i
  }
  wave_lds_sync();
  // X[r][j] = (delta_rj - sum_{m<r} L[r][m] X[m][j]) / L_rr for lane j, L[r][m] = U[m][r]; four partial sums
  double x[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    double p0 = r == i ? 1.0 : 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
    for (int m = 0; m < r; m += 4) {
      p0 -= sT[16 * m + r] * x[m];
      if (m + 1 < r) p1 -= sT[16 * (m + 1) + r] * x[m + 1];
      if (m + 2 < r) p2 -= sT[16 * (m + 2) + r] * x[m + 2];
      if (m + 3 < r) p3 -= sT[16 * (m + 3) + r] * x[m + 3];
    }
    x[r] = ((p0 + p1) + (p2 + p3)) * rdiag[r];
  }
  if (lane < 16) {
#pragma unroll
    for (int q = 0; q < 16; ++q) xt[16 * lane + q] = x[q];  // X^T[lane][q]
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) t[v] = sT[64 * v + lane];
  return ok;
}

// Sum over the 16 lanes of each row of 16 (DPP within rows; every lane gets its row's sum).
__device__ __forceinline__ double dpp_f64(double x, int ctrl_sel) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  switch (ctrl_sel) {
    case 0:
      lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
      hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
      break;
    case 1:
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
      break;
    case 2:
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x124, 0xF, 0xF, false);  // row_ror:4
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x124, 0xF, 0xF, false);
      break;
    default:
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x128, 0xF, 0xF, false);  // row_ror:8
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x128, 0xF, 0xF, false);
      break;
  }
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double sum_row16_dpp(double x) {
  x += dpp_f64(x, 0);
  x += dpp_f64(x, 1);
  x += dpp_f64(x, 2);
  x += dpp_f64(x, 3);
  return x;
}

}  // namespace sba_tiles
