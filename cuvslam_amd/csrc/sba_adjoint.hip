// Implicit gradient through a converged SBA solve (FP64), on the device.
//
// At a solution x* of min_x F(x; theta), F = sum_o rho(r_o^T L_o r_o), r_o = xy_o - pi(x), the gradient
// g(x*, theta) = grad_x F vanishes. For a loss of the solution, dL/dtheta = -a^T dg/dtheta with H a = dL/dx,
// H the full Hessian of F at x* (rho'' terms and the second-order terms of the projection and of the
// left-multiplied twist included). Tangent coordinates: left twists (omega, v) of the free key frames, then
// the free points, as the solver's own step. The observation parameters are xy_o and the symmetric L_o; a
// caller maps them to pixels and intrinsics (x = (u - c_x)/f_x, L_xy = diag(f) L_uv diag(f),
// libs/camera/observation.cpp:57-67).
//
// H a = g is solved by preconditioned conjugate gradients. The preconditioner is the Gauss-Newton system
// 2 sum rho' J^T L J of the converged state, solved through its Schur complement with the reduced camera
// system factored once in FP64 (the solver's own k_schur / k_reduce / k_factor at lambda = 0).
// Kernels: k_adj_obs (per observation: Hessian product, the two Schur passes of the preconditioner, and
// the final per-observation gradients), k_adj_chol (triangular solves with the stored factor), k_adj_vec
// (conjugate-gradient vector updates in one workgroup). No reference counterpart (the reference has no
// differentiation); its analytic Jacobians (schur_complement_bundler_cpu.cpp:101-118) are the anchors.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "sba_device.h"
#include "sba_tiles.h"

using namespace sba_dev;

namespace {

constexpr double kHither64 = static_cast<double>(0.1f);  // epipolar/near_plane.h:33-38

struct Adj {
  const double* spts;   // FP64 state of window 0: points [np][3]
  const double* sposes; // FP64 state of window 0: key frames [nc][12], rows of [R | t]
  const double* v;   // input vector (tangent), for the Hessian product / first Schur pass / gradients
  double* y;         // output vector (Hessian product), or the point accumulator of the second Schur pass
  double* ec;        // pose part of the Schur right-hand side (first pass)
  const double* zc;  // pose part of the Schur solution (second pass)
  double* dxy;       // per-observation gradients, original order
  double* dinfo;
  const int32_t* obs_orig;
  int32_t* done;     // conjugate gradients converged: remaining launches return at once
};

__device__ __forceinline__ double drho64(int kind, double s, double d) {
  if (kind == 0) return (d + 1.0) / (d + s);
  if (d <= 0.0 || s < d * d) return 0.5;
  return 0.5 * d / sqrt(s);
}

__device__ __forceinline__ double d2rho64(int kind, double s, double d) {
  if (kind == 0) return -(d + 1.0) / ((d + s) * (d + s));
  if (d <= 0.0 || s < d * d) return 0.0;
  return -0.25 * d / (s * sqrt(s));
}

struct Lin64 {
  double r[2], J[2][9], L[3], w, w2, q0[3], P[3], Rcr[3][3], Rrw[3][3], D1[3][9];
};

// FP64 linearization of one observation at the FP64 state: p_r = T_rw p, p_c = T_cr p_r, J = dproj D1 with
// D1 = R_cr [-[p_r]x, I, R_rw] (the tangent of a left twist of T_rw and of the point).
__device__ bool lin64(const double* pose, const float4* cam, const double* X, float2 xy, float4 inf, int kind,
                      double delta, Lin64& o) {
  const float4 cr[3] = {cam[0], cam[1], cam[2]};
  for (int a = 0; a < 3; ++a) {
    const double* row = pose + 4 * a;
    const double crow[4] = {cr[a].x, cr[a].y, cr[a].z, cr[a].w};
    for (int k = 0; k < 3; ++k) {
      o.Rrw[a][k] = row[k];
      o.Rcr[a][k] = crow[k];
    }
    o.q0[a] = row[0] * X[0] + row[1] * X[1] + row[2] * X[2] + row[3];
  }
  const double tc[3] = {cr[0].w, cr[1].w, cr[2].w};
  for (int a = 0; a < 3; ++a) o.P[a] = o.Rcr[a][0] * o.q0[0] + o.Rcr[a][1] * o.q0[1] + o.Rcr[a][2] * o.q0[2] + tc[a];
  if (!(o.P[2] > kHither64)) return false;
  const double iz = 1.0 / o.P[2], u = o.P[0] * iz, v = o.P[1] * iz;
  o.r[0] = static_cast<double>(xy.x) - u;
  o.r[1] = static_cast<double>(xy.y) - v;
  const double mq[3][3] = {{0, o.q0[2], -o.q0[1]}, {-o.q0[2], 0, o.q0[0]}, {o.q0[1], -o.q0[0], 0}};
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) {
      double s1 = 0, s3 = 0;
      for (int m = 0; m < 3; ++m) {
        s1 += o.Rcr[a][m] * mq[m][k];
        s3 += o.Rcr[a][m] * o.Rrw[m][k];
      }
      o.D1[a][k] = s1;
      o.D1[a][3 + k] = o.Rcr[a][k];
      o.D1[a][6 + k] = s3;
    }
  for (int k = 0; k < 9; ++k) {
    o.J[0][k] = iz * (o.D1[0][k] - u * o.D1[2][k]);
    o.J[1][k] = iz * (o.D1[1][k] - v * o.D1[2][k]);
  }
  o.L[0] = inf.x;
  o.L[1] = inf.y;
  o.L[2] = inf.z;
  const double Lr0 = o.L[0] * o.r[0] + o.L[1] * o.r[1], Lr1 = o.L[1] * o.r[0] + o.L[2] * o.r[1];
  const double s = o.r[0] * Lr0 + o.r[1] * Lr1;
  o.w = drho64(kind, s, delta);
  o.w2 = d2rho64(kind, s, delta);
  return true;
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// h = H_o v for one observation (9 tangent entries: pose twist, point).
__device__ void hess_obs(const Lin64& L, const double v[9], double h[9]) {
  double u[2] = {0, 0};
  for (int k = 0; k < 9; ++k) {
    u[0] += L.J[0][k] * v[k];
    u[1] += L.J[1][k] * v[k];
  }
  const double Lu[2] = {L.L[0] * u[0] + L.L[1] * u[1], L.L[1] * u[0] + L.L[2] * u[1]};
  const double Lr[2] = {L.L[0] * L.r[0] + L.L[1] * L.r[1], L.L[1] * L.r[0] + L.L[2] * L.r[1]};
  const double rLu = L.r[0] * Lu[0] + L.r[1] * Lu[1];
  for (int k = 0; k < 9; ++k) {
    const double JLr = L.J[0][k] * Lr[0] + L.J[1][k] * Lr[1];
    h[k] = 2 * L.w * (L.J[0][k] * Lu[0] + L.J[1][k] * Lu[1]) + 4 * L.w2 * JLr * rLu;
  }
  // - 2 Hess(c^T pi) v, c = rho' L r: D1^T Hpi D1 v plus the second-order terms of p_r along v
  const double c[2] = {L.w * Lr[0], L.w * Lr[1]};
  const double iz = 1.0 / L.P[2];
  const double g3[3] = {c[0] * iz, c[1] * iz, -(c[0] * L.P[0] + c[1] * L.P[1]) * iz * iz};
  double dv[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 9; ++k) dv[a] += L.D1[a][k] * v[k];
  const double hd[3] = {-c[0] * iz * iz * dv[2], -c[1] * iz * iz * dv[2],
                        -c[0] * iz * iz * dv[0] - c[1] * iz * iz * dv[1] +
                            2 * (c[0] * L.P[0] + c[1] * L.P[1]) * iz * iz * iz * dv[2]};
  double hpsi[9];
  for (int k = 0; k < 9; ++k) hpsi[k] = L.D1[0][k] * hd[0] + L.D1[1][k] * hd[1] + L.D1[2][k] * hd[2];
  double gp[3];
  for (int m = 0; m < 3; ++m) gp[m] = L.Rcr[0][m] * g3[0] + L.Rcr[1][m] * g3[1] + L.Rcr[2][m] * g3[2];
  double Rdp[3];
  for (int m = 0; m < 3; ++m) Rdp[m] = L.Rrw[m][0] * v[6] + L.Rrw[m][1] * v[7] + L.Rrw[m][2] * v[8];
  double wxq[3], avec[3], gxw[3], t1[3], t2[3];
  cross3(v, L.q0, wxq);
  for (int m = 0; m < 3; ++m) avec[m] = 0.5 * wxq[m] + 0.5 * v[3 + m] + Rdp[m];
  cross3(avec, gp, t1);
  cross3(gp, v, gxw);
  cross3(L.q0, gxw, t2);
  for (int m = 0; m < 3; ++m) {
    hpsi[m] += t1[m] + 0.5 * t2[m];
    hpsi[3 + m] += 0.5 * gxw[m];
    hpsi[6 + m] += L.Rrw[0][m] * gxw[0] + L.Rrw[1][m] * gxw[1] + L.Rrw[2][m] * gxw[2];
  }
  for (int k = 0; k < 9; ++k) h[k] -= 2 * hpsi[k];
}

enum AdjMode { kHvp = 0, kSchurRhs = 1, kSchurBack = 2, kObsGrad = 3, kGrad = 4 };

}  // namespace

// One thread per (point, key frame) pair of window 0 over the back-substitution tiles.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_adj_obs(Batch b, Adj A) {
  __shared__ double sPose[6 * kMaxFreePoses];
  const Chunk ch = b.bchunks[blockIdx.x];
  const int w = b.btiles[ch.tile_begin].window;
  if (w != 0 || (MODE != kObsGrad && MODE != kGrad && *A.done)) return;
  const WindowDesc W = b.win[0];
  const int tid = threadIdx.x;
  const int ncf = W.ncf;
  const int pf_end = W.npf;  // window 0 starts at point 0
  const int n6 = 6 * ncf;
  const double delta = W.robust;
  for (int k = tid; k < 6 * ncf; k += kBlock) sPose[k] = 0.0;
  __syncthreads();
  for (int t = ch.tile_begin; t < ch.tile_end; ++t) {
    const Tile T = b.btiles[t];
    const int p0 = b.point_pair[T.point_begin];
    const int npair = b.point_pair[T.point_end] - p0;
    if (tid >= npair) continue;
    const int pair = p0 + tid;
    // the pair's point: the last point of the tile whose first pair is <= pair (binary search)
    int lo = T.point_begin, hi = T.point_end;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (b.point_pair[mid] <= pair) lo = mid; else hi = mid;
    }
    const int i = lo;
    const int j = b.pair_pose[pair];
    const bool pf = i < pf_end, cf = j < ncf;
    const int4 rec = b.pair_rec[pair];
    for (int o = rec.x; o < rec.y; ++o) {
      const float2 xy = b.obs_xy[o];
      const float4 inf = b.obs_info[o];
      const int cam = __float_as_int(inf.w);
      Lin64 L;
      const bool valid = lin64(A.sposes + 12 * j, b.cams + 3 * cam, A.spts + 3 * i, xy, inf, b.loss, delta, L);
      if (MODE == kObsGrad) {
        const int oo = A.obs_orig[o];
        double* gx = A.dxy + 2 * oo;
        double* gi = A.dinfo + 3 * oo;
        gx[0] = gx[1] = gi[0] = gi[1] = gi[2] = 0.0;
        if (!valid) continue;
        double av[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (cf)
          for (int k = 0; k < 6; ++k) av[k] = A.v[6 * j + k];
        if (pf)
          for (int k = 0; k < 3; ++k) av[6 + k] = A.v[n6 + 3 * i + k];
        double Ja[2] = {0, 0};
        for (int k = 0; k < 9; ++k) {
          Ja[0] += L.J[0][k] * av[k];
          Ja[1] += L.J[1][k] * av[k];
        }
        const double LJa[2] = {L.L[0] * Ja[0] + L.L[1] * Ja[1], L.L[1] * Ja[0] + L.L[2] * Ja[1]};
        const double Lr[2] = {L.L[0] * L.r[0] + L.L[1] * L.r[1], L.L[1] * L.r[0] + L.L[2] * L.r[1]};
        const double rLJa = L.r[0] * LJa[0] + L.r[1] * LJa[1];
        gx[0] = 2 * (L.w * LJa[0] + 2 * L.w2 * rLJa * Lr[0]);
        gx[1] = 2 * (L.w * LJa[1] + 2 * L.w2 * rLJa * Lr[1]);
        const double G00 = 2 * (L.w * Ja[0] * L.r[0] + L.w2 * rLJa * L.r[0] * L.r[0]);
        const double G01 = 2 * (L.w * Ja[0] * L.r[1] + L.w2 * rLJa * L.r[0] * L.r[1]);
        const double G10 = 2 * (L.w * Ja[1] * L.r[0] + L.w2 * rLJa * L.r[1] * L.r[0]);
        const double G11 = 2 * (L.w * Ja[1] * L.r[1] + L.w2 * rLJa * L.r[1] * L.r[1]);
        gi[0] = G00;
        gi[1] = G01 + G10;
        gi[2] = G11;
        continue;
      }
      if (!valid) continue;
      if (MODE == kGrad) {
        // grad F = -2 sum rho' J^T L r
        const double Lr[2] = {L.L[0] * L.r[0] + L.L[1] * L.r[1], L.L[1] * L.r[0] + L.L[2] * L.r[1]};
        double h[9];
        for (int k = 0; k < 9; ++k) h[k] = -2 * L.w * (L.J[0][k] * Lr[0] + L.J[1][k] * Lr[1]);
        if (cf)
          for (int k = 0; k < 6; ++k) atomicAdd(&sPose[6 * j + k], h[k]);
        if (pf)
          for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&A.y[n6 + 3 * i + k], h[6 + k]);
      } else if (MODE == kHvp) {
        double v9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (cf)
          for (int k = 0; k < 6; ++k) v9[k] = A.v[6 * j + k];
        if (pf)
          for (int k = 0; k < 3; ++k) v9[6 + k] = A.v[n6 + 3 * i + k];
        double h[9];
        hess_obs(L, v9, h);
        if (cf)
          for (int k = 0; k < 6; ++k) atomicAdd(&sPose[6 * j + k], h[k]);
        if (pf)
          for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&A.y[n6 + 3 * i + k], h[6 + k]);
      } else if (cf && pf) {
        // W = w J_c^T L J_p (the LM system's off-diagonal block; the preconditioner is twice that system)
        double LJp[2][3];
        for (int k = 0; k < 3; ++k) {
          LJp[0][k] = L.w * (L.L[0] * L.J[0][6 + k] + L.L[1] * L.J[1][6 + k]);
          LJp[1][k] = L.w * (L.L[1] * L.J[0][6 + k] + L.L[2] * L.J[1][6 + k]);
        }
        if (MODE == kSchurRhs) {
          // ec_j -= W_ij t_i with t_i = V*^-1 q_p,i
          const float4 a0 = b.paux[3 * i], a1 = b.paux[3 * i + 1];
          const double Vi[3][3] = {{a0.x, a0.y, a0.z}, {a0.y, a0.w, a1.x}, {a0.z, a1.x, a1.y}};
          const double* q = A.v + n6 + 3 * i;
          double tv[3];
          for (int a = 0; a < 3; ++a) tv[a] = Vi[a][0] * q[0] + Vi[a][1] * q[1] + Vi[a][2] * q[2];
          double Lt[2] = {0, 0};
          for (int k = 0; k < 3; ++k) {
            Lt[0] += LJp[0][k] * tv[k];
            Lt[1] += LJp[1][k] * tv[k];
          }
          for (int k = 0; k < 6; ++k) atomicAdd(&sPose[6 * j + k], -(L.J[0][k] * Lt[0] + L.J[1][k] * Lt[1]));
        } else {
          // acc_i += W_ij^T zc_j
          const double* z = A.zc + 6 * j;
          double Jz[2] = {0, 0};
          for (int k = 0; k < 6; ++k) {
            Jz[0] += L.J[0][k] * z[k];
            Jz[1] += L.J[1][k] * z[k];
          }
          for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&A.y[n6 + 3 * i + k], LJp[0][k] * Jz[0] + LJp[1][k] * Jz[1]);
        }
      }
    }
  }
  if (MODE == kHvp || MODE == kSchurRhs || MODE == kGrad) {
    __syncthreads();
    double* dst = MODE == kSchurRhs ? A.ec : A.y;
    for (int k = tid; k < 6 * ncf; k += kBlock)
      if (sPose[k] != 0.0) unsafeAtomicAdd(&dst[k], sPose[k]);
  }
}

namespace {
__device__ __forceinline__ int64_t pk2(int i, int j) { return static_cast<int64_t>(i) * (i + 1) / 2 + j; }
}  // namespace

// L^-1 of the stored FP64 factor (packed lower L, then 1 / L_ii), dense row-major n x n (zero above the diagonal),
// one wavefront per column j: x = L^-1 e_j by forward substitution, each row's dot product over the lanes with a
// fixed butterfly. Run once per factored solution; the preconditioner then applies S^-1 = L^-T L^-1 as two
// parallel products instead of two serial triangular solves.
__global__ __launch_bounds__(64) void k_adj_linv(const double* Lst, int n, double* Linv) {
  __shared__ double x[256];
  const int j = blockIdx.x, lane = threadIdx.x;
  const double* idg = Lst + static_cast<int64_t>(n) * (n + 1) / 2;
  for (int i = lane; i < n; i += 64) x[i] = 0.0;
  sba_tiles::wave_lds_sync();
  for (int i = j; i < n; ++i) {
    double sum = 0.0;
    for (int k = j + lane; k < i; k += 64) sum += Lst[pk2(i, k)] * x[k];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) x[i] = ((i == j ? 1.0 : 0.0) - sum) * idg[i];
    sba_tiles::wave_lds_sync();
  }
  for (int i = lane; i < n; i += 64) Linv[static_cast<int64_t>(i) * n + j] = i >= j ? x[i] : 0.0;
}

// zc = S^-1 ec = L^-T (L^-1 ec) with the dense inverse factor: y = L^-1 ec, one wavefront per row (lanes over
// the row, a fixed butterfly), then zc = L^-T y, one thread per entry.
__global__ __launch_bounds__(256) void k_adj_chol(const double* Linv, int n, const double* ec, double* zc,
                                                  const int32_t* done) {
  __shared__ double y[256];
  if (*done) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  for (int i = wave; i < n; i += nw) {
    const double* row = Linv + static_cast<int64_t>(i) * n;
    double sum = 0.0;
    for (int k = lane; k <= i; k += 64) sum += row[k] * ec[k];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) y[i] = sum;
  }
  __syncthreads();
  for (int jj = tid; jj < n; jj += blockDim.x) {
    double sum = 0.0;
#pragma unroll 4
    for (int i = jj; i < n; ++i) sum += Linv[static_cast<int64_t>(i) * n + jj] * y[i];
    zc[jj] = sum;
  }
}

// Conjugate-gradient vector algebra in one workgroup (nx = 6 ncf + 3 npf entries).
//   op 0: x = 0, r = g, q_p = r_p, ec = r_c                  (set up the first preconditioner call)
//   op 1: z_c = zc / 2, z_p = V*^-1 (q_p - acc) / 2         (finish a preconditioner call; q = r)
//   op 2: first step: p = z, rz = r.z, gg = g.g; y = 0
//   op 3: after y = H p: alpha = rz / p.y, x += alpha p, r -= alpha y, converged?; ec = r_c, acc = 0
//   op 4: after z = M^-1 r: beta = r.z / rz, p = z + beta p, rz = r.z; y = 0
struct Cg {
  double *x, *r, *z, *p, *y, *ec, *zc, *acc, *g, *sc;  // sc: rz, gg, rr
  int32_t* done;
  double tol;
};

__global__ __launch_bounds__(1024) void k_adj_vec(Batch b, Cg c, int nx, int n6, int op) {
  __shared__ double red[2][16];
  const int tid = threadIdx.x;
  if (op != 0 && *c.done) return;
  double s0 = 0.0, s1 = 0.0;
  // the dot products of ops 2-4 read distinct vectors: restrict-qualified, unrolled, so that a thread's loads of
  // several entries are in flight together
  const double* __restrict__ vr = c.r;
  const double* __restrict__ vz = c.z;
  const double* __restrict__ vp = c.p;
  const double* __restrict__ vy = c.y;
  if (op == 3) {
#pragma unroll 8
    for (int k = tid; k < nx; k += 1024) s0 += vp[k] * vy[k];
  } else if (op == 4) {
#pragma unroll 8
    for (int k = tid; k < nx; k += 1024) s0 += vr[k] * vz[k];
  }
  const int gk = blockIdx.x * blockDim.x + tid, gstride = gridDim.x * blockDim.x;
  for (int k = gk; k < nx && op <= 2; k += gstride) {
    if (op == 0) {
      c.x[k] = 0.0;
      c.r[k] = c.g[k];
      if (k < n6) c.ec[k] = c.g[k];
      c.acc[k] = 0.0;
    } else if (op == 1) {
      if (k < n6) {
        c.z[k] = 0.5 * c.zc[k];
      } else {
        const int i = (k - n6) / 3, a = (k - n6) % 3;
        const float4 a0 = b.paux[3 * i], a1 = b.paux[3 * i + 1];
        const double Vi[3][3] = {{a0.x, a0.y, a0.z}, {a0.y, a0.w, a1.x}, {a0.z, a1.x, a1.y}};
        const int base = n6 + 3 * i;
        double s = 0.0;
        for (int m = 0; m < 3; ++m) s += Vi[a][m] * (c.r[base + m] - c.acc[base + m]);
        c.z[k] = 0.5 * s;
      }
    } else {
      c.p[k] = c.z[k];
      s0 += c.r[k] * c.z[k];
      s1 += c.g[k] * c.g[k];
      c.y[k] = 0.0;
    }
  }
  if (op == 0 || op == 1) return;
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = s0;
    red[1][tid >> 6] = s1;
  }
  __syncthreads();
  double t0 = 0.0, t1 = 0.0;
  for (int k = 0; k < 16; ++k) {
    t0 += red[0][k];
    t1 += red[1][k];
  }
  if (op == 2) {
    if (tid == 0) {
      c.sc[0] = t0;
      c.sc[1] = t1;
      c.sc[3] = 0.0;
      if (t1 == 0.0) *c.done = 1;
    }
    return;
  }
  if (op == 3) {
    // negative or zero curvature along p: the Hessian is not positive definite here; stop with the current x
    // (truncated conjugate gradients, as in Steihaug's method)
    if (!(t0 > 0.0)) {
      if (tid == 0) *c.done = 1;
      return;
    }
    const double alpha = c.sc[0] / t0;
    double rr = 0.0;
    double* __restrict__ wx = c.x;
    double* __restrict__ wr = c.r;
    double* __restrict__ wec = c.ec;
    double* __restrict__ wacc = c.acc;
#pragma unroll 4
    for (int k = tid; k < nx; k += 1024) {
      wx[k] += alpha * vp[k];
      const double rk = wr[k] - alpha * vy[k];
      wr[k] = rk;
      rr += rk * rk;
      if (k < n6) wec[k] = rk;
      wacc[k] = 0.0;
    }
    for (int o = 32; o > 0; o >>= 1) rr += __shfl_xor(rr, o);
    __syncthreads();
    if ((tid & 63) == 0) red[0][tid >> 6] = rr;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int k = 0; k < 16; ++k) t += red[0][k];
      c.sc[2] = t;
      c.sc[3] += 1.0;
      if (sqrt(t / c.sc[1]) < c.tol) *c.done = 1;
    }
    return;
  }
  // op 4
  const double beta = t0 / c.sc[0];
  double* __restrict__ wp = c.p;
  double* __restrict__ wy = c.y;
#pragma unroll 4
  for (int k = tid; k < nx; k += 1024) {
    wp[k] = vz[k] + beta * wp[k];
    wy[k] = 0.0;
  }
  __syncthreads();
  if (tid == 0) c.sc[0] = t0;
}

// FP64 state of window 0 from its FP32 state (op 0; free rotations projected onto the rotation group), or a Newton step x <- x - a applied to it (op 1):
// T_j <- Exp(-a_j) T_j for the free key frames (math/twist.cpp:23-41 in FP64), p_i <- p_i - a_i for the free points.
__global__ __launch_bounds__(256) void k_adj_state(Batch b, double* spts, double* sposes, const double* a, int op) {
  const WindowDesc W = b.win[0];
  const int cur = b.st[0].cur;
  const float4* pts = cur ? b.pts1 : b.pts0;
  const float4* poses = cur ? b.poses1 : b.poses0;
  const int n6 = 6 * W.ncf;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < W.np; i += gridDim.x * blockDim.x) {
    if (op == 0) {
      const float4 q = pts[i];
      spts[3 * i] = q.x;
      spts[3 * i + 1] = q.y;
      spts[3 * i + 2] = q.z;
    } else if (i < W.npf) {
      for (int k = 0; k < 3; ++k) spts[3 * i + k] -= a[n6 + 3 * i + k];
    }
  }
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < W.nc; j += gridDim.x * blockDim.x) {
    double* T = sposes + 12 * j;
    if (op == 0) {
      double X[3][3];
      for (int r = 0; r < 3; ++r) {
        const float4 q = poses[3 * j + r];
        X[r][0] = q.x;
        X[r][1] = q.y;
        X[r][2] = q.z;
        T[4 * r + 3] = q.w;
      }
      // the FP32 rotations of the free key frames are orthonormal only to FP32 rounding; the Newton steps move
      // along Exp(xi) R and would otherwise settle on that orbit instead of the rotation group
      if (j < W.ncf) polar3(X);
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) T[4 * r + q] = X[r][q];
    } else if (j < W.ncf) {
      const double w0 = -a[6 * j], w1 = -a[6 * j + 1], w2 = -a[6 * j + 2];
      const double v[3] = {-a[6 * j + 3], -a[6 * j + 4], -a[6 * j + 5]};
      const double theta = fmax(sqrt(w0 * w0 + w1 * w1 + w2 * w2), static_cast<double>(1.1920928955078125e-7f));
      const double c = cos(theta), st = sin(theta) / theta;
      const double n[3] = {w0 / theta, w1 / theta, w2 / theta};
      const double K[3][3] = {{c, -w2 * st, w1 * st}, {w2 * st, c, -w0 * st}, {-w1 * st, w0 * st, c}};
      double E[3][3];
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) E[r][q] = K[r][q] + (1.0 - c) * n[r] * n[q];
      const double nxv[3] = {n[1] * v[2] - n[2] * v[1], n[2] * v[0] - n[0] * v[2], n[0] * v[1] - n[1] * v[0]};
      const double nnv[3] = {n[1] * nxv[2] - n[2] * nxv[1], n[2] * nxv[0] - n[0] * nxv[2], n[0] * nxv[1] - n[1] * nxv[0]};
      double et[3];
      for (int r = 0; r < 3; ++r) et[r] = v[r] + ((1.0 - c) / theta * nxv[r] + (1.0 - st) * nnv[r]);
      double R[3][4];
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 4; ++q) R[r][q] = T[4 * r + q];
      for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) T[4 * r + q] = E[r][0] * R[0][q] + E[r][1] * R[1][q] + E[r][2] * R[2][q];
        T[4 * r + 3] = E[r][0] * R[0][3] + E[r][1] * R[1][3] + E[r][2] * R[2][3] + et[r];
      }
    }
  }
}

namespace sba_dev {

hipError_t launch_adj_state(const Batch& b, double* spts, double* sposes, const double* a, int op, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_state, dim3(64), dim3(256), 0, s, b, spts, sposes, a, op);
  return hipGetLastError();
}

hipError_t launch_adj_obs(const Batch& b, int mode, const double* spts, const double* sposes, const double* v, double* y,
                          double* ec, const double* zc, double* dxy, double* dinfo, const int32_t* orig, int32_t* done,
                          hipStream_t s) {
  Adj A{spts, sposes, v, y, ec, zc, dxy, dinfo, orig, done};
  switch (mode) {
    case kGrad: hipLaunchKernelGGL(k_adj_obs<kGrad>, dim3(b.n_bchunks), dim3(kBlock), 0, s, b, A); break;
    case kHvp: hipLaunchKernelGGL(k_adj_obs<kHvp>, dim3(b.n_bchunks), dim3(kBlock), 0, s, b, A); break;
    case kSchurRhs: hipLaunchKernelGGL(k_adj_obs<kSchurRhs>, dim3(b.n_bchunks), dim3(kBlock), 0, s, b, A); break;
    case kSchurBack: hipLaunchKernelGGL(k_adj_obs<kSchurBack>, dim3(b.n_bchunks), dim3(kBlock), 0, s, b, A); break;
    default: hipLaunchKernelGGL(k_adj_obs<kObsGrad>, dim3(b.n_bchunks), dim3(kBlock), 0, s, b, A); break;
  }
  return hipGetLastError();
}

hipError_t launch_adj_linv(const double* Lst, int n, double* Linv, hipStream_t s) {
  if (n > 256) return hipErrorInvalidValue;
  if (n > 0) hipLaunchKernelGGL(k_adj_linv, dim3(n), dim3(64), 0, s, Lst, n, Linv);
  return hipGetLastError();
}

hipError_t launch_adj_chol(const double* Linv, int n, const double* ec, double* zc, const int32_t* done,
                           hipStream_t s) {
  if (n > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_adj_chol, dim3(1), dim3(256), 0, s, Linv, n, ec, zc, done);
  return hipGetLastError();
}

hipError_t launch_adj_vec(const Batch& b, double* x, double* r, double* z, double* p, double* y, double* ec, double* zc,
                          double* acc, double* g, double* sc, int32_t* done, double tol, int nx, int n6, int op,
                          hipStream_t s) {
  Cg c{x, r, z, p, y, ec, zc, acc, g, sc, done, tol};
  // ops 0 and 1 are elementwise (any number of workgroups); ops 2-4 reduce in one workgroup
  if (op <= 1)
    hipLaunchKernelGGL(k_adj_vec, dim3((nx + 255) / 256), dim3(256), 0, s, b, c, nx, n6, op);
  else
    hipLaunchKernelGGL(k_adj_vec, dim3(1), dim3(1024), 0, s, b, c, nx, n6, op);
  return hipGetLastError();
}

}  // namespace sba_dev
