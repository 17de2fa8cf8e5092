// Drop-in SchurComplementBundlerGpu (libs/sba/schur_complement_bundler_gpu.h:29-38) backed by the MI355X
// SBA through the C ABI of cuvslam_sba.h. solve() has the reference's contract: the problem is updated in
// place; false for an empty problem, an infeasible start (every observation inside the near plane) or a
// device error; true otherwise, also when no step was accepted. No exception escapes.
#pragma once

#include <memory>

#include "sba/bundle_adjustment_problem.h"

struct sba_handle;

namespace cuvslam::sba {

class SchurComplementBundlerGpu {
public:
  SchurComplementBundlerGpu();
  // loss: 0 = Student-t (the reference's), 1 = Huber; device: HIP device ordinal
  explicit SchurComplementBundlerGpu(int loss, int device = 0);
  ~SchurComplementBundlerGpu();
  SchurComplementBundlerGpu(const SchurComplementBundlerGpu&) = delete;
  SchurComplementBundlerGpu& operator=(const SchurComplementBundlerGpu&) = delete;
  bool solve(BundleAdjustmentProblem& problem);

private:
  class Impl;
  std::unique_ptr<Impl> impl;
};

}  // namespace cuvslam::sba
