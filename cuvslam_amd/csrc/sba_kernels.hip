// HIP kernels of the MI355X sliding-window SBA (gfx950).
//
// One Levenberg-Marquardt iteration of every window of a batch is five launches:
//   k_schur     per chunk of a window's points: linearize each (point, key frame) pair (residual, 2x6 pose and
//               2x3 point Jacobians, robust weight), accumulate V_i, e_b, W_ij, U_j, e_a, damp and
//               pseudo-invert V_i, and accumulate the Schur pair products W_ij V_i^-1 W_ik^T on the matrix cores;
//               each chunk writes its partial of the reduced camera system S once, as a record
//   k_reduce    per window: the chunk records summed in a fixed order
//               (reference phases: calc_jacobians_kernel sba_v1.cu:361-460, build_full_system_1/2 :486-631,
//               reduced_system_stage_1/12/2/3 :1016-1204; CPU twin schur_complement_bundler_cpu.cpp:64-287)
//   k_factor    per window: assemble S + lambda diag(U) in FP64, blocked right-looking Cholesky with the
//               forward substitution fused, backward substitution, then Exp(dc) * T re-orthonormalized in
//               FP64 as the trial key frames (reference: cusolverDnSpotrf/Spotrs sba.cpp:606-640,
//               calc_pose_update_kernel sba_v1.cu:823-829, UpdateState cpu.cpp:367-383)
//   k_backsub   per chunk: dp_i = V_i^-1 (e_b - sum_j W_ij^T dc_j), the trial points, the trial cost in FP64,
//               and the predicted-reduction terms dT H d and dT diag(H) d accumulated per observation
//               (reference: calc_point_update_kernel sba_v1.cu:749-786, evaluate_cost_kernel :633-712,
//               predict_reduction sba.cpp:688-727)
//   k_decide    per window: the LM accept/reject decision, damping schedule and stop tests on the device
//               (reference host loop schur_complement_bundler_gpu.cpp:175-253)
// Nothing returns to the host between iterations, so the iterations of a batch replay as one hipGraph.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "sba_device.h"
#include "sba_tiles.h"

using namespace sba_dev;
using namespace sba_tiles;

namespace {

constexpr float kHither = 0.1f;                  // epipolar/near_plane.h:33-38

constexpr float kSqrtFltEps = 3.4526698300124393e-4f;

__device__ __forceinline__ float drho(int kind, float s, float d) {
  // math/robust_cost_function.cpp:62-95 (derivative w.r.t. the squared residual)
  if (kind == 0) return (d + 1.f) / (d + s);
  if (d <= 0.f) return 0.5f;
  if (s < d * d) return 0.5f;
  return 0.5f * d * rsqrtf(s);
}

__device__ __forceinline__ double rho64(int kind, double s, double d) {
  if (kind == 0) return (d + 1.0) * log1p(s / d);
  if (d <= 0.0) return 0.5 * s;
  if (s < d * d) return 0.5 * s;
  return d * sqrt(s) - 0.5 * d * d;
}


// A pose in shared memory: 12 floats, rows of [R | t].
struct Lin {
  float r0, r1, w;
  float jc[2][6];
  float jp[2][3];
};

// Residual, Jacobians and robust weight of one observation (cpu.cpp:84-125). J_c is with respect to a
// left-multiplied twist (omega, v) of rig_from_world, computed as dproj * R_cr * [-[p_r]x, I]: the same
// matrix as the reference's dproj * [-[p_c]x, I] * Adjoint(camera_from_rig).
__device__ __forceinline__ bool linearize(const float* P, const float* C, float px, float py, float pz, float ox,
                                          float oy, float la, float lb, float lc, int kind, float delta, Lin& L) {
  const float prx = P[0] * px + P[1] * py + P[2] * pz + P[3];
  const float pry = P[4] * px + P[5] * py + P[6] * pz + P[7];
  const float prz = P[8] * px + P[9] * py + P[10] * pz + P[11];
  const float pcx = C[0] * prx + C[1] * pry + C[2] * prz + C[3];
  const float pcy = C[4] * prx + C[5] * pry + C[6] * prz + C[7];
  const float pcz = C[8] * prx + C[9] * pry + C[10] * prz + C[11];
  if (!(pcz > kHither)) return false;
  const float iz = 1.f / pcz;
  const float u = pcx * iz, v = pcy * iz;
  L.r0 = ox - u;
  L.r1 = oy - v;
  // A = dproj * R_cr (2x3)
  float A[2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A[0][k] = iz * (C[k] - u * C[8 + k]);
    A[1][k] = iz * (C[4 + k] - v * C[8 + k]);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    // row a of A * (-[p_r]x) = p_r x A_a
    L.jc[a][0] = pry * A[a][2] - prz * A[a][1];
    L.jc[a][1] = prz * A[a][0] - prx * A[a][2];
    L.jc[a][2] = prx * A[a][1] - pry * A[a][0];
    L.jc[a][3] = A[a][0];
    L.jc[a][4] = A[a][1];
    L.jc[a][5] = A[a][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) L.jp[a][k] = A[a][0] * P[k] + A[a][1] * P[4 + k] + A[a][2] * P[8 + k];
  }
  const float s = L.r0 * (la * L.r0 + lb * L.r1) + L.r1 * (lb * L.r0 + lc * L.r1);
  L.w = drho(kind, s, delta);
  return true;
}

// Cyclic Jacobi on a symmetric 3x3 (Golub & Van Loan, Algorithm 8.5.3); a becomes diagonal, v holds the
// eigenvectors as columns.
__device__ __forceinline__ void jrot(float a[3][3], float v[3][3], int p, int q) {
  const float apq = a[p][q];
  if (apq == 0.f) return;
  const float tau = (a[q][q] - a[p][p]) / (2.f * apq);
  const float t = (tau >= 0.f ? 1.f : -1.f) / (fabsf(tau) + sqrtf(1.f + tau * tau));
  const float c = 1.f / sqrtf(1.f + t * t);
  const float s = t * c;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float akp = a[k][p], akq = a[k][q];
    a[k][p] = c * akp - s * akq;
    a[k][q] = s * akp + c * akq;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float apk = a[p][k], aqk = a[q][k];
    a[p][k] = c * apk - s * aqk;
    a[q][k] = s * apk + c * aqk;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float vkp = v[k][p], vkq = v[k][q];
    v[k][p] = c * vkp - s * vkq;
    v[k][q] = s * vkp + c * vkq;
  }
}

// Symmetric factor G of the point-block pseudo-inverse V*^+ = G G^T with the reference's rank rule
// (cpu.cpp:218-262, sba_v1.cu:1016-1069): eigenvalues above 1e-6 of the largest count; rank 3 inverts,
// rank 2 drops the smallest direction, rank <= 1 gives zero.
//
// Well-conditioned blocks take a short path: for a symmetric positive definite 3x3 with eigenvalues
// l1 >= l2 >= l3, the sum of its principal 2x2 minors is m2 = l1 l2 + l1 l3 + l2 l3 >= l1 l2, so
// l3 = det / (l1 l2) >= det / m2, and l1 <= trace; det > 4e-6 trace m2 therefore puts every eigenvalue above
// the rank threshold (with a factor 4 to spare for rounding). The bound is close to l3 itself (m2 ~ l1 l2
// when l3 is small), so only blocks near the threshold take the eigen-decomposition. Then
// V*^-1 = L^-T L^-1 with the Cholesky factor L, and G = L^-T has the same G G^T as the eigen-decomposition
// (only the rank decision needs the eigenvalues).
__device__ __forceinline__ void point_factor(const float V[6], float lambda, float G[3][3]) {
  float a[3][3] = {{V[0] * (1.f + lambda), V[1], V[2]}, {V[1], V[3] * (1.f + lambda), V[4]},
                   {V[2], V[4], V[5] * (1.f + lambda)}};
  {
    const float tr = a[0][0] + a[1][1] + a[2][2];
    const float m00 = a[1][1] * a[2][2] - a[1][2] * a[1][2];
    const float m11 = a[0][0] * a[2][2] - a[0][2] * a[0][2];
    const float m22 = a[0][0] * a[1][1] - a[0][1] * a[0][1];
    const float det = a[0][0] * m00 - a[0][1] * (a[0][1] * a[2][2] - a[0][2] * a[1][2]) +
                      a[0][2] * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
    const float m2 = m00 + m11 + m22;
    if (tr > 0.f && m2 > 0.f && det > 4e-6f * tr * m2 && a[0][0] > 0.f) {
      const float l00 = sqrtf(a[0][0]), i00 = 1.f / l00;
      const float l10 = a[0][1] * i00, l20 = a[0][2] * i00;
      const float d11 = a[1][1] - l10 * l10;
      if (d11 > 0.f) {
        const float l11 = sqrtf(d11), i11 = 1.f / l11;
        const float l21 = (a[1][2] - l20 * l10) * i11;
        const float d22 = a[2][2] - l20 * l20 - l21 * l21;
        if (d22 > 0.f) {
          const float i22 = 1.f / sqrtf(d22);
          // M = L^-1 (lower triangular); G = M^T
          const float m10 = -l10 * i00 * i11, m21 = -l21 * i11 * i22, m20 = (l10 * l21 * i11 - l20) * i00 * i22;
          G[0][0] = i00; G[0][1] = m10; G[0][2] = m20;
          G[1][0] = 0.f; G[1][1] = i11; G[1][2] = m21;
          G[2][0] = 0.f; G[2][1] = 0.f; G[2][2] = i22;
          return;
        }
      }
    }
  }
  float v[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
  const float fro = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2] +
                    2.f * (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2]);
  for (int sweep = 0; sweep < 8; ++sweep) {
    const float off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off <= 1e-14f * fro) break;
    jrot(a, v, 0, 1);
    jrot(a, v, 0, 2);
    jrot(a, v, 1, 2);
  }
  const float e[3] = {a[0][0], a[1][1], a[2][2]};
  const float smax = fmaxf(fabsf(e[0]), fmaxf(fabsf(e[1]), fabsf(e[2])));
  int rank = 0;
  float sc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool keep = smax > 0.f && fabsf(e[k]) > 1e-6f * smax && e[k] > 0.f;
    rank += keep;
    sc[k] = keep ? rsqrtf(e[k]) : 0.f;
  }
  if (rank < 2) sc[0] = sc[1] = sc[2] = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) G[i][k] = v[i][k] * sc[k];
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Sum over the four lanes of a quad, (x0 + x1) + (x2 + x3) in every lane of it (the same bits in all four: the
// two partial sums are added in either order).
__device__ __forceinline__ float quad_sum(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));  // lane ^ 1
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));  // lane ^ 2
  return x;
}

__device__ __forceinline__ void load_pose_lds(float* dst, const float4* src, int n3, int tid, int nthreads) {
  for (int k = tid; k < n3; k += nthreads) {
    const float4 q = src[k];
    dst[4 * k + 0] = q.x;
    dst[4 * k + 1] = q.y;
    dst[4 * k + 2] = q.z;
    dst[4 * k + 3] = q.w;
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// Linearize + Schur complement. One workgroup per chunk (consecutive tiles of one window's points).
//
// The Schur products S -= sum_i W_i V_i*^-1 W_i^T are written as S -= sum_i Z_i Z_i^T with Z_i = W_i G_i and
// G_i G_i^T the damped point-block pseudo-inverse: a symmetric rank-3 update per point. Over the points of a
// tile the Z blocks are staged in shared memory as dense columns (6 rows per free key frame, 3 columns per
// point, zero where a key frame does not see the point), and the chunk's sum Z Z^T accumulates in
// registers on the FP32 matrix cores (v_mfma_f32_16x16x4_f32) over the upper 16x16 tiles of S. Each wave owns
// a fixed quarter of the tiles. Nothing is atomically added in global memory: a chunk writes its record once
// and k_reduce sums the records of a window in a fixed order.
// ---------------------------------------------------------------------------------------------------------
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));


// Upper tile T (row-major over the NT x NT upper triangle) -> (I, J).
constexpr int tri_i(int nt, int T) {
  int I = 0;
  while (T >= nt - I) {
    T -= nt - I;
    ++I;
  }
  return I;
}
constexpr int tri_j(int nt, int T) { return tri_i(nt, T) + (T - (tri_i(nt, T) * nt - tri_i(nt, T) * (tri_i(nt, T) - 1) / 2)); }

__device__ __forceinline__ int upper_tile(int I, int J, int nt) { return I * nt - (I * (I - 1)) / 2 + (J - I); }

template <int NT>
constexpr int tiles_per_wave() { return (NT * (NT + 1) / 2 + 3) / 4; }

// The matrix-core part of one tile of points: acc(I, J) += Zc[16 I.., k] Zc[16 J.., k]^T over its k-steps, for
// the upper tiles T = WAVE, WAVE + 4, ... of the window's NT x NT tile grid. Lane layout of
// v_mfma_f32_16x16x4_f32: A lane l holds A[l % 16][l / 16], B lane l holds B[l / 16][l % 16], D lane l holds
// D[4 (l / 16) + v][l % 16]. NT and the tile coordinates are template parameters: every register index is a
// compile-time constant and each k-step is straight-line code (NT loads, then this wave's instructions).
template <int NT, int WAVE, int t>
__device__ __forceinline__ void mfma_one(const float (&f)[NT], floatx4 (&acc)[tiles_per_wave<NT>()]) {
  constexpr int T = WAVE + 4 * t;
  if constexpr (T < NT * (NT + 1) / 2) {
    constexpr int I = tri_i(NT, T), J = tri_j(NT, T);
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(f[I], f[J], acc[t], 0, 0, 0);
  }
}

template <int NT, int WAVE, int... ts>
__device__ __forceinline__ void mfma_all(const float (&f)[NT], floatx4 (&acc)[tiles_per_wave<NT>()],
                                         std::integer_sequence<int, ts...>) {
  (mfma_one<NT, WAVE, ts>(f, acc), ...);
}

template <int NT, int WAVE>
__device__ __forceinline__ void mfma_tile(const float* sZc, int nslab, int lane, floatx4 (&acc)[tiles_per_wave<NT>()]) {
  const float* col = sZc + (lane >> 4) * kZcLd + (lane & 15);
  for (int s = 0; s < nslab; ++s, col += 4 * kZcLd) {
    float f[NT];
#pragma unroll
    for (int I = 0; I < NT; ++I) f[I] = col[16 * I];
    mfma_all<NT, WAVE>(f, acc, std::make_integer_sequence<int, tiles_per_wave<NT>()>{});
  }
}

template <int NT, int WAVE, int t>
__device__ __forceinline__ void store_one(float* rec, int lane, const floatx4 (&acc)[tiles_per_wave<NT>()]) {
  constexpr int T = WAVE + 4 * t;
  if constexpr (T < NT * (NT + 1) / 2) {
    constexpr int I = tri_i(NT, T), J = tri_j(NT, T);
    float* dst = rec + 256 * upper_tile(I, J, NT);
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[64 * v + lane] = acc[t][v];
  }
}

// Mapped windows: the record holds the upper 6 x 6 key-frame blocks (la <= lb) one after another, 36 floats each
// (row-major; of a diagonal block only the upper triangle is written), so that a key-frame block of the window's
// system reads each chunk's part of it as one contiguous piece.
template <int NT, int WAVE, int t>
__device__ __forceinline__ void store_one_blocked(float* rec, int lane, int nloc,
                                                  const floatx4 (&acc)[tiles_per_wave<NT>()]) {
  constexpr int T = WAVE + 4 * t;
  if constexpr (T < NT * (NT + 1) / 2) {
    constexpr int I = tri_i(NT, T), J = tri_j(NT, T);
    const int cl = 16 * J + (lane & 15), lb = cl / 6;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rl = 16 * I + 4 * (lane >> 4) + v, la = rl / 6;
      if (rl <= cl && cl < 6 * nloc)
        rec[36 * (la * nloc - (la * (la - 1)) / 2 + (lb - la)) + 6 * (rl - 6 * la) + (cl - 6 * lb)] = acc[t][v];
    }
  }
}

template <int NT, int WAVE, int... ts>
__device__ __forceinline__ void store_all_blocked(float* rec, int lane, int nloc,
                                                  const floatx4 (&acc)[tiles_per_wave<NT>()],
                                                  std::integer_sequence<int, ts...>) {
  (store_one_blocked<NT, WAVE, ts>(rec, lane, nloc, acc), ...);
}

template <int NT, int WAVE, int... ts>
__device__ __forceinline__ void store_all(float* rec, int lane, const floatx4 (&acc)[tiles_per_wave<NT>()],
                                          std::integer_sequence<int, ts...>) {
  (store_one<NT, WAVE, ts>(rec, lane, acc), ...);
}

template <int NT, int WAVE>
__device__ __forceinline__ void store_tiles(float* rec, int lane, int nloc, bool blocked,
                                            const floatx4 (&acc)[tiles_per_wave<NT>()]) {
  if (blocked) store_all_blocked<NT, WAVE>(rec, lane, nloc, acc, std::make_integer_sequence<int, tiles_per_wave<NT>()>{});
  else store_all<NT, WAVE>(rec, lane, acc, std::make_integer_sequence<int, tiles_per_wave<NT>()>{});
}

// Workgroup barrier for shared memory only: waits for this wave's LDS operations, not for its global loads, so
// loads issued for the next tile stay in flight across it. The empty asm keeps the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt left at their maxima (gfx9 encoding)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// A pair's record and first observation, loaded one tile ahead.
struct PairPre {
  int4 rec;  // (first observation, observation end, key frame, point index in the tile)
  float2 xy;
  float4 inf;
};

// Per-pair contributions are reduced without floating-point atomics: each pair thread stores its U/e_a (27) and
// V/e_b (9) terms as columns of a shared scratch, and fixed owner threads sum them. A key frame's terms in a tile
// are found through a (tile point, key frame) -> pair table; a point's pairs are contiguous.
constexpr int kUVals = 27;               // U (21 upper entries) and e (6) of one key frame
constexpr int kUGroup = 5;               // values per owner thread
constexpr int kUGroups = (kUVals + kUGroup - 1) / kUGroup;  // 6
static_assert(kUGroups * kMaxFreePoses <= kBlock - kTilePoints, "owner threads of U and V must not overlap");
constexpr int kScr = kTilePairs + 1;     // scratch row stride: owners reading one pair's values hit distinct banks
static_assert((kUVals + 9) * kScr <= kKt * kZcLd, "the scratch aliases the staged Z columns");

// One chunk of tiles of one window with NT = ceil(6 N_c,free / 16) tiles per side of its reduced system.
//
// Threads: pair thread e < npair handles pair p0 + e (linearize, W, Z = W G); U owner thread
// tid < 6 ncf owns values [5 g, 5 g + 5) of key frame j = tid / 6 (g = tid % 6) for the whole chunk, in
// registers; V owner thread kBlock - 16 + tp sums point tp's V, e_b and factors its block. Per tile:
//   (b) linearize; scratch <- per-pair U/e_a and V/e_b terms; pair table <- pair index
//   (c) owners sum the scratch (U into registers; V -> G, G^T e_b into sAcc and paux)
//   (z) zero the staged Z columns (they alias the scratch)
//   (d) Z = W G into the staged columns; e_c terms -Z G^T e_b into their own scratch
//   (e) e owners sum the e_c terms; matrix cores: acc += Zc Zc^T
// POSES_GLOBAL: the key frames are read from global memory (cached) instead of the shared copy, for batches whose
// largest window would otherwise leave room for only one workgroup per compute unit.
template <int NT, bool POSES_GLOBAL>
__device__ __forceinline__ void schur_chunk(const Batch& b, const Chunk& ch, const WindowDesc& W, float lam, int cur,
                                            float* sm) {
  constexpr int nt = NT;
  const float4* __restrict__ pts = cur ? b.pts1 : b.pts0;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int ncf = W.ncf;
  const int nloc = ch.nloc;  // free key frames staged by this chunk (local index jl; = ncf unless mapped)
  const float* sPose = POSES_GLOBAL ? reinterpret_cast<const float*>((cur ? b.poses1 : b.poses0) + 3 * W.pose_off)
                                    : sm + b.lds_pose;
  float* sCam = sm + b.lds_cam;
  float* sAcc = sm + b.lds_acc;  // 12 per tile point: G (9) and G^T e_b (3)
  int* sEnt = reinterpret_cast<int*>(sm + b.lds_ent);  // [tile point][free key frame]: (tile << 9) | pair
  float* sEc = sm + b.lds_ec;    // [6][kScr] e_c terms
  float* sW = sm + b.lds_w;      // [18][kScr] W_ij of each pair (6x3, row-major)
  float4* sPts = reinterpret_cast<float4*>(sm + b.lds_pts);  // [2][kTilePoints] the tile's points, double-buffered
  float* sZc = sm + b.lds_zc;    // Z columns: [k][row], k = 3 * tile point + c, row = 6 * key frame + r
  float* sScrU = sZc;            // [27][kScr] per-pair U/e_a terms (aliases the Z columns)
  float* sScrV = sZc + kUVals * kScr;  // [9][kScr] per-pair V/e_b terms
  const int kind = b.loss;
  const float delta = W.robust;
  const int pf_end = W.point_off + W.npf;
  // U owner: key frame uj, values [u0, u1)
  const bool u_owner = tid < kUGroups * nloc;
  const int uj = tid / kUGroups, u0 = kUGroup * (tid % kUGroups), u1 = min(u0 + kUGroup, kUVals);
  float accU[kUGroup] = {0.f, 0.f, 0.f, 0.f, 0.f};
  // V owners of tile point vp: the last wavefront, four lanes per point (lane vs takes every fourth pair) when
  // the U owners leave it free; otherwise its last 16 lanes, one per point. The choice depends on the window
  // only (a mapped window always takes one lane), so a point's sums do not depend on the chunking.
  const bool v4 = !W.mapped && kUGroups * nloc <= kBlock - 64;
  const int vp = v4 ? (tid - (kBlock - 64)) >> 2 : tid - (kBlock - kTilePoints);
  const int vs = v4 ? tid & 3 : 0, vstep = v4 ? 4 : 1;

  for (int k = tid; k < kTilePoints * kMaxFreePoses; k += kBlock) sEnt[k] = -1;
  // the first tile's pair data, points and point pair ranges
  Tile Tn = b.tiles[ch.tile_begin];
  int pp_a = 0, pp_e = 0;  // pair range of point vp, loaded one tile ahead
  if (vp >= 0 && vp < Tn.point_end - Tn.point_begin) {
    pp_a = b.point_pair[Tn.point_begin + vp] + Tn.pair_shift;
    pp_e = b.point_pair[Tn.point_begin + vp + 1] + Tn.pair_shift;
  }
  PairPre pre{};
  if (tid < Tn.pair_end - Tn.pair_begin) {
    pre.rec = b.pair_rec[Tn.pair_begin + tid];
    pre.xy = b.pair_xy[Tn.pair_begin + tid];
    pre.inf = b.pair_info[Tn.pair_begin + tid];
  }
  if (tid < Tn.point_end - Tn.point_begin) sPts[(ch.tile_begin & 1) * kTilePoints + tid] = pts[Tn.point_begin + tid];

  floatx4 acc[tiles_per_wave<NT>()];
#pragma unroll
  for (int t = 0; t < tiles_per_wave<NT>(); ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int t = ch.tile_begin; t < ch.tile_end; ++t) {
    const Tile T = Tn;
    if (t + 1 < ch.tile_end) Tn = b.tiles[t + 1];  // used by the prefetch in (e)
    const int p0 = T.pair_begin;
    const int npair = T.pair_end - p0;
    const int npts = T.point_end - T.point_begin;
    const int kpad = (3 * npts + 3) & ~3;
    const int stamp = (t - ch.tile_begin) << 9;
    lds_barrier();  // the previous tile's shared arrays are free
    // (b) one thread per pair: linearize its observations (one per camera), store W and the U and V terms
    int tp = 0, j = 0, jl = 0;
    bool pair_pf = false, pair_cf = false;
    if (tid < npair) {
      tp = pre.rec.w & 255;
      jl = pre.rec.w >> 8;
      const int i = T.point_begin + tp;
      j = pre.rec.z & 0xffff;
      pair_pf = i < pf_end;
      pair_cf = j < ncf && jl != 255;  // a free key frame staged by this chunk
      const float4 p = sPts[(t & 1) * kTilePoints + tp];
      // the pair's terms go straight to its own scratch column (no other thread touches it): the first valid
      // observation stores, later ones (other cameras of a rig) add; a pair without one stores zeros
      // (store mode is a template argument, so that the common case of one observation per pair only writes)
      auto obs = [&](int o, auto store) -> bool {
        auto put = [&](float* arr, int k, float v) {
          float* q = arr + k * kScr + tid;
          if constexpr (decltype(store)::value) *q = v;
          else *q += v;
        };
        const float2 xy = o == pre.rec.x ? pre.xy : b.obs_xy[o];
        const float4 inf = o == pre.rec.x ? pre.inf : b.obs_info[o];
        const int cam = __float_as_int(inf.w);
        Lin L;
        if (!linearize(sPose + 12 * j, sCam + 12 * cam, p.x, p.y, p.z, xy.x, xy.y, inf.x, inf.y, inf.z, kind, delta, L))
          return false;
        float LJc[2][6], LJp[2][3];
        const float wr0 = L.w * (inf.x * L.r0 + inf.y * L.r1);
        const float wr1 = L.w * (inf.y * L.r0 + inf.z * L.r1);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          LJc[0][k] = L.w * (inf.x * L.jc[0][k] + inf.y * L.jc[1][k]);
          LJc[1][k] = L.w * (inf.y * L.jc[0][k] + inf.z * L.jc[1][k]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          LJp[0][k] = L.w * (inf.x * L.jp[0][k] + inf.y * L.jp[1][k]);
          LJp[1][k] = L.w * (inf.y * L.jp[0][k] + inf.z * L.jp[1][k]);
        }
        if (pair_cf) {
          int u = 0;
#pragma unroll
          for (int r = 0; r < 6; ++r) {
#pragma unroll
            for (int c = r; c < 6; ++c) put(sScrU, u++, L.jc[0][r] * LJc[0][c] + L.jc[1][r] * LJc[1][c]);
            put(sScrU, 21 + r, L.jc[0][r] * wr0 + L.jc[1][r] * wr1);
          }
          if (pair_pf) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
              for (int c = 0; c < 3; ++c) put(sW, 3 * r + c, L.jc[0][r] * LJp[0][c] + L.jc[1][r] * LJp[1][c]);
          }
        }
        if (pair_pf) {
          put(sScrV, 0, L.jp[0][0] * LJp[0][0] + L.jp[1][0] * LJp[1][0]);
          put(sScrV, 1, L.jp[0][0] * LJp[0][1] + L.jp[1][0] * LJp[1][1]);
          put(sScrV, 2, L.jp[0][0] * LJp[0][2] + L.jp[1][0] * LJp[1][2]);
          put(sScrV, 3, L.jp[0][1] * LJp[0][1] + L.jp[1][1] * LJp[1][1]);
          put(sScrV, 4, L.jp[0][1] * LJp[0][2] + L.jp[1][1] * LJp[1][2]);
          put(sScrV, 5, L.jp[0][2] * LJp[0][2] + L.jp[1][2] * LJp[1][2]);
#pragma unroll
          for (int k = 0; k < 3; ++k) put(sScrV, 6 + k, L.jp[0][k] * wr0 + L.jp[1][k] * wr1);
        }
        return true;
      };
      bool first = true;
      int o = pre.rec.x;
      for (; first && o < pre.rec.y; ++o) first = !obs(o, std::true_type{});
      for (; o < pre.rec.y; ++o) obs(o, std::false_type{});
      if (first) {
        if (pair_cf) {
          for (int k = 0; k < kUVals; ++k) sScrU[k * kScr + tid] = 0.f;
          if (pair_pf)
            for (int k = 0; k < 18; ++k) sW[k * kScr + tid] = 0.f;
        }
        if (pair_pf)
          for (int k = 0; k < 9; ++k) sScrV[k * kScr + tid] = 0.f;
      }
      if (pair_cf) sEnt[tp * kMaxFreePoses + jl] = stamp | tid;
    }
    lds_barrier();
    // (c) owners sum the pair terms: U and e_a of key frame uj into registers; V, e_b of point vp, then its
    // damped pseudo-inverse V*^+ = G G^T
    if (u_owner) {
      // four tile points at a time: their 20 values and the next four points' table entries, each batch of loads
      // in flight together (values past u1 are read and dropped: rows u0 + v <= 29 lie inside the scratch, whose
      // 9 V rows follow the 27 U rows; entries past the tile's points are read and not used)
      int en_next[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) en_next[j] = sEnt[j * kMaxFreePoses + uj];
#pragma unroll 1
      for (int q0 = 0; q0 < npts; q0 += 4) {
        int en[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) en[j] = en_next[j];
        if (q0 + 4 < kTilePoints) {
#pragma unroll
          for (int j = 0; j < 4; ++j) en_next[j] = sEnt[(q0 + 4 + j) * kMaxFreePoses + uj];
        }
        float x[4][kUGroup];
        bool ok[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ok[j] = q0 + j < npts && en[j] >= 0 && (en[j] & ~511) == stamp;
          const int col = ok[j] ? (en[j] & 511) : 0;
#pragma unroll
          for (int v = 0; v < kUGroup; ++v) x[j][v] = sScrU[(u0 + v) * kScr + col];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int v = 0; v < kUGroup; ++v) accU[v] += ok[j] && u0 + v < u1 ? x[j][v] : 0.f;
      }
    }
    if (vp >= 0 && vp < npts && T.point_begin + vp < pf_end) {
      const int i = T.point_begin + vp;
      const int a = pp_a - p0, e = pp_e - p0;
      // pairs a + vs, a + vs + vstep, ... two at a time (18 loads in flight), then the four lanes of a point add
      // their sums
      float V[9] = {};
      for (int k = a + vs; k < e; k += 2 * vstep) {
        const int k2 = min(k + vstep, e - 1);
        const bool two = k + vstep < e;
        float x[9], y[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
          x[m] = sScrV[m * kScr + k];
          y[m] = sScrV[m * kScr + k2];
        }
#pragma unroll
        for (int m = 0; m < 9; ++m) V[m] += two ? x[m] + y[m] : x[m];
      }
      if (v4) {
#pragma unroll
        for (int m = 0; m < 9; ++m) V[m] = quad_sum(V[m]);
      }
      const float eb[3] = {V[6], V[7], V[8]};
      float G[3][3];
      point_factor(V, lam, G);
      float tv[3], Vi[6], gb[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) tv[k] = G[0][k] * eb[0] + G[1][k] * eb[1] + G[2][k] * eb[2];
      Vi[0] = G[0][0] * G[0][0] + G[0][1] * G[0][1] + G[0][2] * G[0][2];
      Vi[1] = G[0][0] * G[1][0] + G[0][1] * G[1][1] + G[0][2] * G[1][2];
      Vi[2] = G[0][0] * G[2][0] + G[0][1] * G[2][1] + G[0][2] * G[2][2];
      Vi[3] = G[1][0] * G[1][0] + G[1][1] * G[1][1] + G[1][2] * G[1][2];
      Vi[4] = G[1][0] * G[2][0] + G[1][1] * G[2][1] + G[1][2] * G[2][2];
      Vi[5] = G[2][0] * G[2][0] + G[2][1] * G[2][1] + G[2][2] * G[2][2];
#pragma unroll
      for (int k = 0; k < 3; ++k) gb[k] = G[k][0] * tv[0] + G[k][1] * tv[1] + G[k][2] * tv[2];
      if (vs == 0) {
        float4* ax = b.paux + 3 * i;
        ax[0] = make_float4(Vi[0], Vi[1], Vi[2], Vi[3]);
        ax[1] = make_float4(Vi[4], Vi[5], gb[0], gb[1]);
        ax[2] = make_float4(gb[2], V[0], V[3], V[5]);
        float* ac = sAcc + 12 * vp;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) ac[3 * r + c] = G[r][c];
        ac[9] = tv[0];
        ac[10] = tv[1];
        ac[11] = tv[2];
      }
    }
    lds_barrier();
    // (z) zero the staged columns that the matrix cores read: kpad k-rows x 16 nt rows, 16-byte stores
    {
      const int q = 4 * nt;
      float4* z4 = reinterpret_cast<float4*>(sZc);
      for (int k = tid; k < kpad * q; k += kBlock) z4[(k / q) * (kZcLd / 4) + k % q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    lds_barrier();
    // (d) one thread per (free point, free key frame) pair: Z = W G into the staged columns, e_c terms -Z G^T e_b
    if (tid < npair && pair_cf) {
      float ec[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (pair_pf) {
        const float* a = sAcc + 12 * tp;
        float* zc = sZc + 3 * tp * kZcLd + 6 * jl;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          float zr[3];
          const float w0 = sW[(3 * r) * kScr + tid], w1 = sW[(3 * r + 1) * kScr + tid],
                      w2 = sW[(3 * r + 2) * kScr + tid];
#pragma unroll
          for (int c = 0; c < 3; ++c) zr[c] = w0 * a[c] + w1 * a[3 + c] + w2 * a[6 + c];
#pragma unroll
          for (int c = 0; c < 3; ++c) zc[c * kZcLd + r] = zr[c];
          ec[r] = -(zr[0] * a[9] + zr[1] * a[10] + zr[2] * a[11]);
        }
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) sEc[r * kScr + tid] = ec[r];
    }
    lds_barrier();
    // (e) the next tile's pair data is requested; matrix cores: acc += Zc Zc^T; the e owners sum the e_c terms;
    // the next tile's points go to the other half of sPts (read after the next barrier)
    const bool has_next = t + 1 < ch.tile_end;
    float4 pn = make_float4(0.f, 0.f, 0.f, 0.f);
    const int npts_next = has_next ? Tn.point_end - Tn.point_begin : 0;
    if (has_next && tid < Tn.pair_end - Tn.pair_begin) {
      pre.rec = b.pair_rec[Tn.pair_begin + tid];
      pre.xy = b.pair_xy[Tn.pair_begin + tid];
      pre.inf = b.pair_info[Tn.pair_begin + tid];
    }
    if (tid < npts_next) pn = pts[Tn.point_begin + tid];
    if (vp >= 0 && vp < npts_next) {
      pp_a = b.point_pair[Tn.point_begin + vp] + Tn.pair_shift;
      pp_e = b.point_pair[Tn.point_begin + vp + 1] + Tn.pair_shift;
    }
    switch (tid >> 6) {
      case 0: mfma_tile<NT, 0>(sZc, kpad / 4, lane, acc); break;
      case 1: mfma_tile<NT, 1>(sZc, kpad / 4, lane, acc); break;
      case 2: mfma_tile<NT, 2>(sZc, kpad / 4, lane, acc); break;
      default: mfma_tile<NT, 3>(sZc, kpad / 4, lane, acc); break;
    }
    // after the matrix-core instructions are issued, so that these shared-memory sums overlap their execution
    if (u_owner && u1 > 21) {
      // as in (c), the next table entries loaded with the current values; rows -1 .. 8 around the six e_c rows
      // are read (the pair table lies before them, the W scratch after them) and those outside 21 <= u0 + v < u1
      // dropped
      int en_next[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) en_next[j] = sEnt[j * kMaxFreePoses + uj];
#pragma unroll 1
      for (int q0 = 0; q0 < npts; q0 += 4) {
        int en[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) en[j] = en_next[j];
        if (q0 + 4 < kTilePoints) {
#pragma unroll
          for (int j = 0; j < 4; ++j) en_next[j] = sEnt[(q0 + 4 + j) * kMaxFreePoses + uj];
        }
        float x[4][kUGroup];
        bool ok[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ok[j] = q0 + j < npts && en[j] >= 0 && (en[j] & ~511) == stamp;
          const int col = ok[j] ? (en[j] & 511) : 0;
#pragma unroll
          for (int v = 0; v < kUGroup; ++v) x[j][v] = sEc[(u0 + v - 21) * kScr + col];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int v = 0; v < kUGroup; ++v) accU[v] += ok[j] && u0 + v >= 21 && u0 + v < u1 ? x[j][v] : 0.f;
      }
    }
    if (tid < npts_next) sPts[((t + 1) & 1) * kTilePoints + tid] = pn;
  }
  // the chunk's record: Schur tiles (lane order; key-frame blocks in a mapped window), then U (21 per key frame),
  // e (6), diag U (6)
  float* rec = b.part + ch.part_off;
  const bool blocked = W.mapped != 0;
  if (nt > 0 && ch.nt > 0) switch (tid >> 6) {
    case 0: store_tiles<NT, 0>(rec, lane, nloc, blocked, acc); break;
    case 1: store_tiles<NT, 1>(rec, lane, nloc, blocked, acc); break;
    case 2: store_tiles<NT, 2>(rec, lane, nloc, blocked, acc); break;
    default: store_tiles<NT, 3>(rec, lane, nloc, blocked, acc); break;
  }
  if (u_owner) {
    float* ru = rec + (blocked ? 36 * (nloc * (nloc + 1) / 2) : 256 * (nt * (nt + 1) / 2));
    constexpr int kDiag[6] = {0, 6, 11, 15, 18, 20};  // diagonal entries of the packed upper 6x6
#pragma unroll
    for (int v = 0; v < kUGroup; ++v) {
      const int u = u0 + v;
      if (u >= u1) continue;
      if (u < 21) {
        ru[21 * uj + u] = accU[v];
#pragma unroll
        for (int r = 0; r < 6; ++r)
          if (u == kDiag[r]) ru[27 * nloc + 6 * uj + r] = accU[v];
      } else {
        ru[21 * nloc + 6 * uj + (u - 21)] = accU[v];
      }
    }
  }
}

template <bool POSES_GLOBAL>
__device__ __forceinline__ void schur_dispatch(const Batch& b, const Chunk& ch, const WindowDesc& W, float lam, int cur,
                                               float* sm) {
  static_assert(kNT == 15, "the dispatch below covers 1..15 tiles");
  switch (ch.nt) {
    case 0:  // no free key frame: only the point blocks (V*^-1, V*^-1 e_b) are needed; the record is empty
    case 1: schur_chunk<1, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 2: schur_chunk<2, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 3: schur_chunk<3, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 4: schur_chunk<4, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 5: schur_chunk<5, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 6: schur_chunk<6, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 7: schur_chunk<7, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 8: schur_chunk<8, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 9: schur_chunk<9, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 10: schur_chunk<10, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 11: schur_chunk<11, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 12: schur_chunk<12, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 13: schur_chunk<13, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    case 14: schur_chunk<14, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
    default: schur_chunk<15, POSES_GLOBAL>(b, ch, W, lam, cur, sm); break;
  }
}

}  // namespace

template <bool POSES_GLOBAL>
__global__ __launch_bounds__(kBlock, 2) void k_schur(Batch b) {
  extern __shared__ float sm[];
  const Chunk ch = b.chunks[blockIdx.x];
  const int w = ch.window;
  const WindowDesc W = b.win[w];
  const WindowState* St = b.st + w;
  if (St->done) return;
  const int cur = St->cur;
  const float lam = St->lambda;
  const int tid = threadIdx.x;
  if (!POSES_GLOBAL) load_pose_lds(sm + b.lds_pose, (cur ? b.poses1 : b.poses0) + 3 * W.pose_off, 3 * W.nc, tid, kBlock);
  load_pose_lds(sm + b.lds_cam, b.cams + 3 * W.cam_off, 3 * W.ncam, tid, kBlock);
  schur_dispatch<POSES_GLOBAL>(b, ch, W, lam, cur, sm);
}

// Sums the chunk records of each window (deterministic; no atomics). A window of many chunks: the four
// wavefronts of a workgroup take consecutive quarters of the chunks for the same 64 record entries, each
// summed in chunk order with eight loads in flight, and the quarters are added in a fixed order. A window of
// few chunks: one entry per thread over all its chunks.
__global__ __launch_bounds__(256) void k_reduce(Batch b) {
  __shared__ float sq[3][64];
  // this workgroup's window: the last w with red_first[w] <= blockIdx.x
  int lo = 0, hi = b.n_windows - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.red_first[mid] <= static_cast<int>(blockIdx.x)) lo = mid;
    else hi = mid - 1;
  }
  const int w = lo, blk = blockIdx.x - b.red_first[w];
  const WindowDesc W = b.win[w];
  if (b.st[w].done || W.mapped) return;
  if (W.nchunks < 64) {
    const int64_t e = static_cast<int64_t>(blk) * 256 + threadIdx.x;
    if (e >= W.rec) return;
    const float* src = b.part + W.part_off + e;
    float a0 = 0.f, a1 = 0.f;
    int c = 0;
    for (; c + 2 <= W.nchunks; c += 2) {
      a0 += src[c * static_cast<int64_t>(W.rec)];
      a1 += src[(c + 1) * static_cast<int64_t>(W.rec)];
    }
    if (c < W.nchunks) a0 += src[c * static_cast<int64_t>(W.rec)];
    b.sred[W.sred_off + e] = a0 + a1;
    return;
  }
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t e = static_cast<int64_t>(blk) * 64 + lane;
  const int64_t rec = W.rec;
  const int c0 = (W.nchunks * q) / 4, c1 = (W.nchunks * (q + 1)) / 4;
  float s = 0.f;
  if (e < rec) {
    const float* src = b.part + W.part_off + e;
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int c = c0;
    for (; c + 8 <= c1; c += 8) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += src[(c + k) * rec];
    }
    for (; c < c1; ++c) a[0] += src[c * rec];
    s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
  if (q > 0) sq[q - 1][lane] = s;
  __syncthreads();
  if (q == 0 && e < rec) b.sred[W.sred_off + e] = ((s + sq[0][lane]) + sq[1][lane]) + sq[2][lane];
}

// ---------------------------------------------------------------------------------------------------------
// FP64 factor/solve of the reduced camera system, one workgroup per window.
//
// S + lambda diag(U) (n = 6 x free key frames <= 240) is cut into 16 x 16 tiles; A = U^T U is factored as the
// upper tiles U_JI (J <= I) in the FP64 matrix-core layout (element (4v + lane / 16, lane % 16) in register v
// of the lane, i.e. 64 v + lane in row-major order): wavefronts 1..7 keep the off-diagonal tiles in registers,
// wavefront 0 the diagonal tiles in LDS. Right-looking, per tile column K:
//   panel     U_KI = X_KK A_KI with the inverse X_KK = U_KK^-T (4 MFMA 16x16x4), published to LDS, and the
//             forward substitution e_I -= U_KI^T y_K;
//   trailing  A_JI -= U_KJ^T U_KI for K < J <= I (4 MFMA each); wavefront 0 first updates (K+1, K+1) and
//             factors it (one wavefront, pivots broadcast with readlane, then X_{K+1,K+1} by forward
//             substitution), so the serial 16 x 16 factor overlaps the other wavefronts' updates; y_K = X_KK e_K
//             is formed at the start of the next panel step. Two barriers per tile column.
// Backward substitution U x = y per tile column: x_K = X_KK^T r_K, then r_J -= U_JK x_K by the tiles' owners.
// (Golub & Van Loan, Matrix Computations, 4.2.4 block Cholesky and 3.1 triangular systems.)
// ---------------------------------------------------------------------------------------------------------
namespace {

constexpr int kFT = 15;     // tile rows of the largest reduced system (240 = 15 x 16)
constexpr int kSlots = 15;  // register tiles per wavefront: 7 wavefronts hold the <= 105 off-diagonal tiles

__device__ __forceinline__ int64_t pk(int i, int j) { return static_cast<int64_t>(i) * (i + 1) / 2 + j; }

}  // namespace

__global__ __launch_bounds__(kFactorThreads) void k_factor(Batch b, double* dbg_S, double* store_L) {
  static_assert(kFactorThreads == 512, "wavefront 0 plus 7 wavefronts of off-diagonal tiles");
  __shared__ double vec[256];  // e, updated by the forward substitution
  __shared__ double ys[256];   // y, then r of the backward substitution
  __shared__ double xs[256];   // x
  __shared__ double du[256];   // diag(U)
  __shared__ double idg[256];  // 1 / L_ii
  __shared__ double red[32];
  __shared__ double sT[256];          // the diagonal tile being factored
  __shared__ double sDg[kFT * 256];   // diagonal tiles (wavefront 0), row-major
  __shared__ double sXT[kFT * 256];   // X_KK^T of every tile column
  __shared__ double sP[kFT * 256];    // panel tiles U_KI of the current tile column
  __shared__ int flag;
  const int w = blockIdx.x;
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done || W.mapped) return;  // mapped windows: sba_large.hip
  const int tid = threadIdx.x;
  const int lane = tid & 63, li = lane >> 4, lj = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ncf = W.ncf;
  const int n = 6 * ncf;
  const int T = W.nt;  // tile rows, 16 T >= n
  const float lam = St->lambda;
  const float* R = b.sred + W.sred_off;  // sum Z Z^T upper tiles (matrix-core lane order), then U, e, diag U
  const float* RU = R + 256 * (T * (T + 1) / 2);
  const float* RE = RU + 21 * ncf;
  const float* RD = RE + 6 * ncf;

  // element (4v + lane/16, lane%16) of upper tile (J, I) of A = U - sum Z Z^T + lambda diag(U); rows and columns
  // past n are identity. Stored sum Z Z^T tiles hold element (row, col) at 64 (row & 3) + 16 (row >> 2) + col;
  // diagonal tiles are read above their diagonal.
  auto element = [&](int J, int I, int v) -> double {
    // every load in bounds and no branch around the Z Z^T loads, so that a wavefront's tile loads are in flight
    // together
    const int rl = 4 * v + li, r = 16 * J + rl, c = 16 * I + lj;
    const bool inside = r < n && c < n;
    const bool mir = J == I && rl > lj;
    const int er = mir ? lj : rl, ec = mir ? rl : lj;
    const float zz = R[256 * upper_tile(J, I, T) + 64 * (er & 3) + 16 * (er >> 2) + ec];
    const int lo = min(r, c), hi = max(r, c), pose = lo / 6, rr = lo - 6 * pose, cc = hi - 6 * pose;
    const bool blk = inside && cc < 6;
    // U only inside a key frame's 6 x 6 block, diag U only on the diagonal: the other lanes load nothing
    const float u = blk ? RU[21 * pose + 6 * rr - (rr * (rr - 1)) / 2 + (cc - rr)] : 0.f;
    const float dg = r == c ? RD[min(r, n - 1)] : 0.f;
    const double val = -static_cast<double>(zz) + (blk ? static_cast<double>(u) : 0.0) +
                       (r == c ? static_cast<double>(lam) * static_cast<double>(dg) : 0.0);
    return inside ? val : (r == c ? 1.0 : 0.0);
  };
  auto dbg_tile = [&](int J, int I, int v, double val) {
    const int r = 16 * J + 4 * v + li, c = 16 * I + lj;
    if (r < n && c < n) {
      dbg_S[static_cast<int64_t>(r) * n + c] = val;
      dbg_S[static_cast<int64_t>(c) * n + r] = val;
    }
  };
  for (int i = tid; i < 256; i += kFactorThreads) {
    vec[i] = i < n ? static_cast<double>(RE[i]) : 0.0;
    du[i] = i < n ? static_cast<double>(RD[i]) : 0.0;
  }
  if (tid == 0) flag = 0;
  if (dbg_S && w == 0)
    for (int i = tid; i < n; i += kFactorThreads) dbg_S[static_cast<int64_t>(n) * n + i] = static_cast<double>(RE[i]);
  const bool keep = store_L && w == 0;
  const int64_t npk = static_cast<int64_t>(n) * (n + 1) / 2;
  // the factor (packed lower L = U^T, then 1 / L_ii) kept for later solves with other right-hand sides
  auto keep_tile = [&](int J, int I, int v, double val) {
    const int j = 16 * J + 4 * v + li, i = 16 * I + lj;  // U[j][i] = L[i][j]
    if (i < n && j <= i) store_L[pk(i, j)] = val;
  };

  // the diagonal tiles, kept in LDS, are assembled by all wavefronts
  for (int J = wave; J < T; J += kFactorThreads / 64) {
    double val[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) val[v] = element(J, J, v);
#pragma unroll
    for (int v = 0; v < 4; ++v) sDg[256 * J + 64 * v + lane] = val[v];
    if (dbg_S && w == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) dbg_tile(J, J, v, val[v]);
    }
  }

  if (wave == 0) {
    // wavefront 0: the diagonal factors
    __syncthreads();
    // 2. factorization; K = -1 only factors tile (0, 0). The other wavefronts run their own copy of the loop
    //    with the same barriers.
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
      // the next diagonal tile first: update, factor; then the remaining diagonal tiles
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
    if (keep) {
      for (int J = 0; J < T; ++J)
#pragma unroll
        for (int v = 0; v < 4; ++v) keep_tile(J, J, v, sDg[256 * J + 64 * v + lane]);
    } else if (flag == 0) {
      // 3. backward substitution U x = y, tile column K from T - 1 down: x_K = X_KK^T r_K (here), then
      //    r_J -= U_JK x_K (the owners of (J, K))
      for (int K = T - 1; K >= 0; --K) {
        const double* xt = sXT + 256 * K;
        double p = 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) p += xt[16 * lj + 4 * li + kk] * ys[16 * K + 4 * li + kk];
        p = sum_rows4(p);
        if (lane < 16) xs[16 * K + lane] = p;
        __syncthreads();
      }
    }
  } else {
    // wavefronts 1..7: the off-diagonal tiles (J < I) in registers, tile m of the row-major order in slot
    // (m / 7) of wavefront 1 + m % 7
    int tJ[kSlots], tI[kSlots];
    {
      const int noff = T * (T - 1) / 2;
#pragma unroll
      for (int s = 0; s < kSlots; ++s) {
        int m = 7 * s + wave - 1, J = 0;
        if (m < noff) {
          while (m >= T - 1 - J) {
            m -= T - 1 - J;
            ++J;
          }
          tJ[s] = J;
          tI[s] = J + 1 + m;
        } else {
          tJ[s] = -1;
          tI[s] = -1;
        }
      }
    }
    double4v acc[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      acc[s] = double4v{0.0, 0.0, 0.0, 0.0};
      if (tJ[s] >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[s][v] = element(tJ[s], tI[s], v);
      }
    }
    if (dbg_S && w == 0) {
#pragma unroll
      for (int s = 0; s < kSlots; ++s)
        if (tJ[s] >= 0)
#pragma unroll
          for (int v = 0; v < 4; ++v) dbg_tile(tJ[s], tI[s], v, acc[s][v]);
    }
    __syncthreads();
    for (int K = -1; K < T; ++K) {
      if (flag) break;
      if (K >= 0) {
        // panel U_KI = X_KK A_KI (the B operand of chunk q is register q of the tile itself), published to
        // LDS, and e_I -= U_KI^T y_K
        const double* xt = sXT + 256 * K;
        double yk = 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) yk += xt[16 * (4 * li + kk) + lj] * vec[16 * K + 4 * li + kk];
        yk = sum_rows4(yk);
        double yv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) yv[v] = __shfl(yk, 4 * v + li);  // y_K[4v + lane/16]
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
          if (tJ[s] == K) {
            const int I = in_loop(tI[s]);
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
        // trailing A_JI -= U_KJ^T U_KI, K < J < I; and the diagonal tiles past K + 1 (in LDS), J = K + 2 + m
        // by wavefront 1 + m % 7, while wavefront 0 factors tile K + 1
        for (int J = K + 1 + wave; J < T; J += 7) {
          double* tj = sDg + 256 * J;
          double4v t;
#pragma unroll
          for (int v = 0; v < 4; ++v) t[v] = tj[64 * v + lane];
          t = tile_update(t, sP + 256 * J, sP + 256 * J, lane);
#pragma unroll
          for (int v = 0; v < 4; ++v) tj[64 * v + lane] = t[v];
        }
#pragma unroll
        for (int s = 0; s < kSlots; ++s)
          if (tJ[s] > K) {
            acc[s] = tile_update(acc[s], sP + 256 * in_loop(tJ[s]), sP + 256 * in_loop(tI[s]), lane);
            __builtin_amdgcn_sched_barrier(0);  // one tile's operands live at a time
          }
      }
      __syncthreads();
    }
    if (keep) {
#pragma unroll
      for (int s = 0; s < kSlots; ++s)
        if (tJ[s] >= 0)
#pragma unroll
          for (int v = 0; v < 4; ++v) keep_tile(tJ[s], tI[s], v, acc[s][v]);
    } else if (flag == 0) {
      for (int K = T - 1; K >= 0; --K) {
        const double* xt = sXT + 256 * K;
        double p = 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) p += xt[16 * lj + 4 * li + kk] * ys[16 * K + 4 * li + kk];
        p = sum_rows4(p);  // x_K[lane % 16]
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
          if (tI[s] == K && tJ[s] >= 0) {
            const int J = in_loop(tJ[s]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const double t = sum_row16_dpp(acc[s][v] * p);
              if (lj == 0) ys[16 * J + 4 * v + li] -= t;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  const bool ok = flag == 0;

  if (keep) {
    for (int i = tid; i < n; i += kFactorThreads) store_L[npk + i] = idg[i];
    if (tid == 0) St->chol_ok = ok ? 1 : 0;
    return;
  }

  if (ok) {
    // 4. the pose step, its max-norm and dc^T diag(U) dc
    double sc = 0.0;
    float mx = 0.f;
    for (int i = tid; i < n; i += kFactorThreads) {
      const double x = xs[i];
      b.dc[W.dc_off + i] = static_cast<float>(x);
      b.dc64[W.dc_off + i] = x;
      sc += x * x * du[i];
      mx = fmaxf(mx, fabsf(static_cast<float>(x)));
    }
    sc = wave_sum(sc);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) {
      red[wave] = sc;
      red[16 + wave] = static_cast<double>(mx);
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0, m = 0.0;
      for (int k = 0; k < kFactorThreads / 64; ++k) {
        s += red[k];
        m = fmax(m, red[16 + k]);
      }
      St->scal_pose = s;
      St->maxdc = static_cast<float>(m);
    }
    // 5. trial key frames: Exp(dc_j) * T_j, rotation re-orthonormalized, rounded to FP32
    write_trial_poses(b, W, St->cur, xs, tid, kFactorThreads);
  }
  if (tid == 0) St->chol_ok = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------------------
// Back-substitution, trial cost (FP64) and predicted-reduction terms. INITIAL: cost of the current state.
// ---------------------------------------------------------------------------------------------------------
template <bool INITIAL>
__global__ __launch_bounds__(kBlock) void k_backsub(Batch b) {
  // the window's key frames (current and trial) and pose step: 30 floats per key frame of the batch's largest
  // window, in the dynamic shared memory
  extern __shared__ float sDyn[];
  float* sPose = sDyn;
  float* sTrial = sDyn + 12 * b.max_nc;
  float* sDc = sDyn + 24 * b.max_nc;
  __shared__ float sCam[12 * kMaxCameras];
  __shared__ float sAcc[4 * kBlock];  // per tile point (up to kBlock): dp
  __shared__ float sPair[3 * kBlock];  // per pair: its part of sum_j W_ij^T dc_j
  __shared__ double sRed[3 * (kBlock / 64)];
  __shared__ int sRedI[kBlock / 64];
  __shared__ float sRedM[kBlock / 64];
  const Chunk ch = b.bchunks[blockIdx.x];
  const int w = b.btiles[ch.tile_begin].window;
  const WindowDesc W = b.win[w];
  WindowState* St = b.st + w;
  if (St->done) return;
  if (!INITIAL && !St->chol_ok) return;
  const int cur = St->cur;
  const int tid = threadIdx.x;
  const int ncf = W.ncf;
  const float4* pts = cur ? b.pts1 : b.pts0;
  float4* pts_trial = cur ? b.pts0 : b.pts1;
  load_pose_lds(sPose, (cur ? b.poses1 : b.poses0) + 3 * W.pose_off, 3 * W.nc, tid, kBlock);
  if (!INITIAL) {
    load_pose_lds(sTrial, (cur ? b.poses0 : b.poses1) + 3 * W.pose_off, 3 * W.nc, tid, kBlock);
    for (int k = tid; k < 6 * ncf; k += kBlock) sDc[k] = b.dc[W.dc_off + k];
  }
  load_pose_lds(sCam, b.cams + 3 * W.cam_off, 3 * W.ncam, tid, kBlock);
  const int kind = b.loss;
  const float delta = W.robust;
  const double delta64 = static_cast<double>(W.robust);
  const int pf_end = W.point_off + W.npf;
  double cost = 0.0, hess = 0.0, scal = 0.0;
  int skipped = 0;
  float maxdp = 0.f;
  __syncthreads();  // key frames, cameras and dc in shared memory

  // this thread's pair of the tile: record (first observation, end, key frame, point in the tile) and first
  // observation; the next tile's are requested before this tile's (c), so that their loads overlap it
  Tile Tn = b.btiles[ch.tile_begin];
  int4 rec_n = make_int4(0, 0, 0, 0);
  float2 xy_n = make_float2(0.f, 0.f);
  float4 inf_n = make_float4(0.f, 0.f, 0.f, 0.f);
  if (tid < Tn.pair_end - Tn.pair_begin) {
    rec_n = b.pair_rec[Tn.pair_begin + tid];
    xy_n = b.pair_xy[Tn.pair_begin + tid];
    inf_n = b.pair_info[Tn.pair_begin + tid];
  }
  for (int t = ch.tile_begin; t < ch.tile_end; ++t) {
    const Tile T = Tn;
    if (t + 1 < ch.tile_end) Tn = b.btiles[t + 1];
    const int p0 = T.pair_begin;
    const int npair = T.pair_end - p0;
    const int npts = T.point_end - T.point_begin;
    const int4 rec = tid < npair ? rec_n : make_int4(0, 0, 0, 0);
    const float2 xy0 = xy_n;
    const float4 inf0 = inf_n;
    // No barrier here: this tile's (a) writes only sPair, which the previous tile last read before its (b)-(c)
    // barrier, and (b) writes sAcc after the (a)-(b) barrier, which every thread reaches after its previous (c).
    // the first observation's J_c dc_j, J_p and weight, kept from (a) for the predicted reduction in (c):
    // kept 0 = not linearized in (a), 1 = kept, 2 = behind the camera
    int kept = 0;
    float ku0 = 0.f, ku1 = 0.f, kw = 0.f, kjp[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (!INITIAL) {
      // (a) per pair: sum over its observations of w J_p^T Lambda J_c dc_j, into the pair's own slot (the point's
      // sum over its pairs is taken in pair order in (b): no floating-point atomics)
      if (tid < npair) {
        float acc[3] = {0.f, 0.f, 0.f};
        const int tp = rec.z >> 16;  // point inside this back-substitution tile
        const int i = T.point_begin + tp;
        const int j = rec.z & 0xffff;
        if (i < pf_end && j < ncf) {
          const float4 p = pts[i];
          const float* d = sDc + 6 * j;
          for (int o = rec.x; o < rec.y; ++o) {
            const float2 xy = o == rec.x ? xy0 : b.obs_xy[o];
            const float4 inf = o == rec.x ? inf0 : b.obs_info[o];
            Lin L;
            if (!linearize(sPose + 12 * j, sCam + 12 * __float_as_int(inf.w), p.x, p.y, p.z, xy.x, xy.y, inf.x, inf.y,
                           inf.z, kind, delta, L)) {
              if (o == rec.x) kept = 2;
              continue;
            }
            float u0 = 0.f, u1 = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
              u0 += L.jc[0][k] * d[k];
              u1 += L.jc[1][k] * d[k];
            }
            if (o == rec.x) {
              kept = 1;
              ku0 = u0;
              ku1 = u1;
              kw = L.w;
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                kjp[0][k] = L.jp[0][k];
                kjp[1][k] = L.jp[1][k];
              }
            }
            const float lu0 = L.w * (inf.x * u0 + inf.y * u1), lu1 = L.w * (inf.y * u0 + inf.z * u1);
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += L.jp[0][k] * lu0 + L.jp[1][k] * lu1;
          }
        }
        sPair[3 * tid] = acc[0];
        sPair[3 * tid + 1] = acc[1];
        sPair[3 * tid + 2] = acc[2];
      }
      __syncthreads();
      // (b) one thread per point: dp = V*^-1 e_b - V*^-1 (sum W^T dc); trial point
      if (tid < npts) {
        const int i = T.point_begin + tid;
        const float4 p = pts[i];
        float dp[3] = {0.f, 0.f, 0.f};
        if (i < pf_end) {
          const float4 a0 = b.paux[3 * i], a1 = b.paux[3 * i + 1], a2 = b.paux[3 * i + 2];
          const float Vi[3][3] = {{a0.x, a0.y, a0.z}, {a0.y, a0.w, a1.x}, {a0.z, a1.x, a1.y}};
          // the point's pairs in order, two at a time
          const int qa = b.point_pair[i] - p0, qe = b.point_pair[i + 1] - p0;
          float s[3] = {0.f, 0.f, 0.f};
          for (int q = qa; q < qe; q += 2) {
            const int q2 = min(q + 1, qe - 1);
            const bool two = q + 1 < qe;
            float x[3], y[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              x[k] = sPair[3 * q + k];
              y[k] = sPair[3 * q2 + k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) s[k] += two ? x[k] + y[k] : x[k];
          }
          const float gb[3] = {a1.z, a1.w, a2.x};
#pragma unroll
          for (int k = 0; k < 3; ++k) dp[k] = gb[k] - (Vi[k][0] * s[0] + Vi[k][1] * s[1] + Vi[k][2] * s[2]);
          const double d0 = dp[0], d1 = dp[1], d2 = dp[2];
          scal += d0 * d0 * a2.y + d1 * d1 * a2.z + d2 * d2 * a2.w;
          maxdp = fmaxf(maxdp, fmaxf(fabsf(dp[0]), fmaxf(fabsf(dp[1]), fabsf(dp[2]))));
        }
        sAcc[4 * tid] = dp[0];
        sAcc[4 * tid + 1] = dp[1];
        sAcc[4 * tid + 2] = dp[2];
        pts_trial[i] = make_float4(p.x + dp[0], p.y + dp[1], p.z + dp[2], 0.f);
        if (b.dbg_dp && w == 0) b.dbg_dp[i - W.point_off] = make_float4(dp[0], dp[1], dp[2], 0.f);
      }
      __syncthreads();
    }
    if (t + 1 < ch.tile_end && tid < Tn.pair_end - Tn.pair_begin) {
      rec_n = b.pair_rec[Tn.pair_begin + tid];
      xy_n = b.pair_xy[Tn.pair_begin + tid];
      inf_n = b.pair_info[Tn.pair_begin + tid];
    }
    // (c) one thread per pair: trial cost in FP64, and w |J_c dc + J_p dp|^2_Lambda
    if (tid < npair) {
      const int tp = rec.z >> 16;
      const int i = T.point_begin + tp;
      const int j = rec.z & 0xffff;
      const float4 p = pts[i];
      float dp[3] = {0.f, 0.f, 0.f}, dcj[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (!INITIAL) {
        dp[0] = sAcc[4 * tp];
        dp[1] = sAcc[4 * tp + 1];
        dp[2] = sAcc[4 * tp + 2];
        if (j < ncf)
#pragma unroll
          for (int k = 0; k < 6; ++k) dcj[k] = sDc[6 * j + k];
      }
      const float* Pt = INITIAL ? (sPose + 12 * j) : (sTrial + 12 * j);
      const double qx = static_cast<double>(p.x + dp[0]), qy = static_cast<double>(p.y + dp[1]),
                   qz = static_cast<double>(p.z + dp[2]);
      const double prx = Pt[0] * qx + Pt[1] * qy + Pt[2] * qz + Pt[3];
      const double pry = Pt[4] * qx + Pt[5] * qy + Pt[6] * qz + Pt[7];
      const double prz = Pt[8] * qx + Pt[9] * qy + Pt[10] * qz + Pt[11];
      for (int o = rec.x; o < rec.y; ++o) {
        const float2 xy = o == rec.x ? xy0 : b.obs_xy[o];
        const float4 inf = o == rec.x ? inf0 : b.obs_info[o];
        const float* C = sCam + 12 * __float_as_int(inf.w);
        const double pcx = C[0] * prx + C[1] * pry + C[2] * prz + C[3];
        const double pcy = C[4] * prx + C[5] * pry + C[6] * prz + C[7];
        const double pcz = C[8] * prx + C[9] * pry + C[10] * prz + C[11];
        if (pcz > static_cast<double>(kHither)) {
          const double r0 = static_cast<double>(xy.x) - pcx / pcz, r1 = static_cast<double>(xy.y) - pcy / pcz;
          const double s = r0 * (inf.x * r0 + inf.y * r1) + r1 * (inf.y * r0 + inf.z * r1);
          cost += rho64(kind, s, delta64);
        } else {
          ++skipped;
        }
        if (!INITIAL && o == rec.x && kept != 0) {
          if (kept == 1) {
            float g0 = ku0, g1 = ku1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              g0 += kjp[0][k] * dp[k];
              g1 += kjp[1][k] * dp[k];
            }
            hess += static_cast<double>(kw * (g0 * (inf.x * g0 + inf.y * g1) + g1 * (inf.y * g0 + inf.z * g1)));
          }
        } else if (!INITIAL) {
          Lin L;
          if (linearize(sPose + 12 * j, C, p.x, p.y, p.z, xy.x, xy.y, inf.x, inf.y, inf.z, kind, delta, L)) {
            float g0 = 0.f, g1 = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
              g0 += L.jc[0][k] * dcj[k];
              g1 += L.jc[1][k] * dcj[k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              g0 += L.jp[0][k] * dp[k];
              g1 += L.jp[1][k] * dp[k];
            }
            hess += static_cast<double>(L.w * (g0 * (inf.x * g0 + inf.y * g1) + g1 * (inf.y * g0 + inf.z * g1)));
          }
        }
      }
    }
  }
  // workgroup reduction, then one set of atomics per workgroup
  cost = wave_sum(cost);
  hess = wave_sum(hess);
  scal = wave_sum(scal);
  skipped = wave_sum(skipped);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) maxdp = fmaxf(maxdp, __shfl_xor(maxdp, o));
  const int warp = tid >> 6;
  if ((tid & 63) == 0) {
    sRed[3 * warp] = cost;
    sRed[3 * warp + 1] = hess;
    sRed[3 * warp + 2] = scal;
    sRedI[warp] = skipped;
    sRedM[warp] = maxdp;
  }
  __syncthreads();
  if (tid == 0) {
    double c = 0.0, h = 0.0, s = 0.0;
    int sk = 0;
    float m = 0.f;
    for (int k = 0; k < kBlock / 64; ++k) {
      c += sRed[3 * k];
      h += sRed[3 * k + 1];
      s += sRed[3 * k + 2];
      sk += sRedI[k];
      m = fmaxf(m, sRedM[k]);
    }
    b.bpart[blockIdx.x] = CostPart{c, INITIAL ? 0.0 : h, INITIAL ? 0.0 : s, sk, INITIAL ? 0.f : m};
  }
}

// ---------------------------------------------------------------------------------------------------------
// LM decision per window (schur_complement_bundler_gpu.cpp:147-253). INITIAL: the setup checks (:149-173).
// ---------------------------------------------------------------------------------------------------------
// One wavefront per window: lane l sums the window's chunk slots l, l + 64, ... in order, then a fixed butterfly
// adds the lanes (the same bits on every run); lane 0 decides.
template <bool INITIAL>
__global__ __launch_bounds__(64) void k_decide(Batch b) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;
  const WindowDesc W = b.win[w];
  WindowState& S = b.st[w];
  if (S.done) return;
  if (!INITIAL && !S.chol_ok) {  // no trial state was evaluated
    if (lane == 0) {
      S.iterations += 1;
      TraceEntry* tr = b.trace && S.iterations <= kTrace ? b.trace + static_cast<int64_t>(w) * kTrace + S.iterations - 1
                                                         : nullptr;
      if (tr) *tr = TraceEntry{0.0, 0.0, S.lambda, 2, {0, 0}};
      S.lambda *= 5.f;
      if (S.iterations >= W.max_iterations) S.done = 1;
    }
    return;
  }
  double c = 0.0, h = 0.0, sp = 0.0;
  int sk = 0;
  float mdp = 0.f;
  for (int k = lane; k < W.nbchunks; k += 64) {
    const CostPart q = b.bpart[W.bchunk_begin + k];
    c += q.cost;
    h += q.hess;
    sp += q.scal;
    sk += q.skipped;
    mdp = fmaxf(mdp, q.maxdp);
  }
  c = wave_sum(c);
  h = wave_sum(h);
  sp = wave_sum(sp);
  sk = wave_sum(sk);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mdp = fmaxf(mdp, __shfl_xor(mdp, o));
  if (lane != 0) return;
  const double cost = sk == W.no ? static_cast<double>(INFINITY) : c;
  const double hess = h, scal_pt = sp;
  const float maxdp = mdp;
  if (INITIAL) {
    S.initial_cost = S.current_cost = S.last_cost = cost;
    S.lambda = 0.001f;
    if (cost < static_cast<double>(kFltEps)) {
      S.done = 1;
      S.status = 0;
    } else if (isinf(cost)) {
      S.done = 1;
      S.status = 2;  // SBA_E_INFEASIBLE
    } else if (W.max_iterations <= 0) {
      S.done = 1;
    }
    return;
  }
  S.iterations += 1;
  const bool last = S.iterations >= W.max_iterations;
  TraceEntry* tr = b.trace && S.iterations <= kTrace ? b.trace + static_cast<int64_t>(w) * kTrace + S.iterations - 1
                                                     : nullptr;
  if (tr) *tr = TraceEntry{0.0, cost, S.lambda, 0, {0, 0}};
  if (S.current_cost < S.initial_cost * static_cast<double>(kFltEps)) {
    S.done = 1;
    if (tr) tr->flags = 4;
    return;
  }
  const double lam = static_cast<double>(S.lambda);
  const double pred = (hess + 2.0 * lam * (S.scal_pose + scal_pt)) / S.current_cost;
  if (!b.fixed_iterations && pred < static_cast<double>(kSqrtFltEps) && S.maxdc < kSqrtFltEps && maxdp < kSqrtFltEps) {
    S.done = 1;
    if (tr) tr->flags = 4;
    return;
  }
  const double rho = (1.0 - cost / S.current_cost) / pred;
  if (tr) tr->rho = rho;
  if (rho > 0.25) {
    if (rho > 0.75 && S.lambda * 0.125f > 0.f) S.lambda *= 0.125f;
    S.current_cost = cost;
    S.last_cost = cost;
    S.cur ^= 1;
    S.accepted += 1;
    if (tr) tr->flags = 1;
  } else {
    S.lambda *= 5.f;
  }
  if (last) S.done = 1;
}

// Restores the uploaded initial state and clears every accumulator (the first node of a solve).
__global__ void k_reset(Batch b) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = gid; i < b.n_points; i += stride) b.pts0[i] = b.pts_init[i];
  for (int64_t i = gid; i < 3 * static_cast<int64_t>(b.n_poses); i += stride) b.poses0[i] = b.poses_init[i];
  for (int64_t i = gid; i < b.n_windows; i += stride) {
    WindowState s = {};
    s.lambda = 0.001f;
    b.st[i] = s;
  }
}

// Copies each window's final state (whichever buffer is current) to the output streams.
__global__ void k_gather(Batch b) {
  const int w = blockIdx.y;
  const WindowDesc W = b.win[w];
  const int cur = b.st[w].cur;
  const float4* pts = cur ? b.pts1 : b.pts0;
  const float4* poses = cur ? b.poses1 : b.poses0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < W.np; i += gridDim.x * blockDim.x)
    b.out_pts[W.point_off + i] = pts[W.point_off + i];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * W.nc; i += gridDim.x * blockDim.x)
    b.out_poses[3 * W.pose_off + i] = poses[3 * W.pose_off + i];
}

// One v_mfma_f32_16x16x4_f32 on fixed inputs, to check the lane layout the Schur kernel relies on:
// A (16x4) lane l holds A[l % 16][l / 16]; B (4x16) lane l holds B[l / 16][l % 16];
// D (16x16) lane l holds D[4 (l / 16) + v][l % 16] in its v-th value.
__global__ void k_mfma_layout(const float* A, const float* B, float* D) {
  const int l = threadIdx.x;
  const float a = A[(l % 16) * 4 + l / 16];
  const float b = B[(l / 16) * 16 + l % 16];
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  for (int v = 0; v < 4; ++v) D[(4 * (l / 16) + v) * 16 + l % 16] = acc[v];
}

// Clears a 256-byte aligned device range in 16-byte stores (a kernel rather than a memset node, so that a solve's
// set-up can sit inside a caller's stream capture). The tail below 16 bytes is cleared one byte per thread.
__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, int64_t n16, uint8_t* __restrict__ tail,
                                                   int32_t ntail) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t i = gid; i < n16; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
  if (gid < ntail) tail[gid] = 0;
}

// A plain streaming copy of a known size, the calibration of HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_calibrate(float4* __restrict__ dst, const float4* __restrict__ src, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------------------------
// Launch wrappers (host side)
// ---------------------------------------------------------------------------------------------------------
namespace sba_dev {

hipError_t launch_reset(const Batch& b, hipStream_t s) {
  hipLaunchKernelGGL(k_reset, dim3(1024), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_schur(const Batch& b, hipStream_t s) {
  // key frames staged in shared memory unless the layout reads them from global memory (lds_pose < 0)
  if (b.lds_pose < 0)
    hipLaunchKernelGGL(k_schur<true>, dim3(b.n_chunks), dim3(kBlock), static_cast<size_t>(b.lds_words) * 4, s, b);
  else
    hipLaunchKernelGGL(k_schur<false>, dim3(b.n_chunks), dim3(kBlock), static_cast<size_t>(b.lds_words) * 4, s, b);
  return hipGetLastError();
}

hipError_t launch_reduce(const Batch& b, hipStream_t s) {
  if (b.red_blocks > 0) hipLaunchKernelGGL(k_reduce, dim3(b.red_blocks), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_factor(const Batch& b, double* dbg_S, hipStream_t s, double* store_L) {
  hipLaunchKernelGGL(k_factor, dim3(b.n_windows), dim3(kFactorThreads), 0, s,
                     b, dbg_S, store_L);
  return hipGetLastError();
}

// Dynamic shared memory of the back-substitution: current and trial key frames and the pose step of the batch's
// largest window (the static arrays take about 9 KB besides).
static size_t backsub_lds_bytes(const Batch& b) { return sizeof(float) * 30 * static_cast<size_t>(b.max_nc > 0 ? b.max_nc : 1); }
static_assert(sizeof(float) * 30 * kMaxPoses + 9 * 1024 <= 160 * 1024, "back-substitution shared memory");

hipError_t launch_backsub(const Batch& b, bool initial, hipStream_t s) {
  if (initial)
    hipLaunchKernelGGL(k_backsub<true>, dim3(b.n_bchunks), dim3(kBlock), backsub_lds_bytes(b), s, b);
  else
    hipLaunchKernelGGL(k_backsub<false>, dim3(b.n_bchunks), dim3(kBlock), backsub_lds_bytes(b), s, b);
  return hipGetLastError();
}

hipError_t launch_decide(const Batch& b, bool initial, hipStream_t s) {
  if (initial)
    hipLaunchKernelGGL(k_decide<true>, dim3(b.n_windows), dim3(64), 0, s, b);
  else
    hipLaunchKernelGGL(k_decide<false>, dim3(b.n_windows), dim3(64), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_gather(const Batch& b, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(16, b.n_windows), dim3(256), 0, s, b);
  return hipGetLastError();
}

// Dynamic shared memory up to the 160 KiB of a gfx950 CU needs no attribute on ROCm; this only clears the
// runtime's last-error slot so that a stale error is not reported by the first launch of a handle.
hipError_t launch_copy_calibrate(float4* dst, const float4* src, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_calibrate, dim3(2048), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  const int64_t n16 = static_cast<int64_t>(bytes / 16);
  const int32_t ntail = static_cast<int32_t>(bytes % 16);
  const int64_t blocks = std::min<int64_t>(std::max<int64_t>((n16 + 255) / 256, 1), 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, static_cast<uint4*>(p), n16,
                     static_cast<uint8_t*>(p) + 16 * n16, ntail);
  return hipGetLastError();
}

hipError_t launch_mfma_layout(const float* A, const float* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

hipError_t set_kernel_attributes() {
  (void)hipGetLastError();
  return hipSuccess;
}

}  // namespace sba_dev
