// HIP kernel of the batched mono SBA solve: one workgroup per window runs the whole Levenberg-Marquardt loop of
// MonoSBASolver::solve (mono_sba_solver.cpp:205-301) on the device, from the window's input poses and points to
// its solved state, covariance and statistics (the solver body is sba_mono_body.h).
#include <hip/hip_runtime.h>

#include "sba_mono_body.h"

using namespace sba_mono;

__global__ __launch_bounds__(kMonoThreads) void k_mono_solve(const MonoJob* jobs, int lds_doubles) {
  extern __shared__ double dyn[];  // the eigen-decomposition's matrices when they fit (lds_doubles)
  __shared__ MonoShared sh;
  __shared__ MonoJob job;
  if (threadIdx.x == 0) job = jobs[blockIdx.x];
  __syncthreads();
  MonoSolve s{job.w, sh, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x)};
  s.lds = lds_doubles > 0 ? dyn : nullptr;
  s.lds_cap = lds_doubles;
  s.run();
  s.export_state(job.poses_out, job.pts_out);
}

namespace sba_mono {

hipError_t launch_mono_solve(const MonoJob* jobs, int n, int lds_doubles, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mono_solve, dim3(n), dim3(kMonoThreads), sizeof(double) * static_cast<size_t>(lds_doubles), s,
                     jobs, lds_doubles);
  return hipGetLastError();
}

}  // namespace sba_mono
