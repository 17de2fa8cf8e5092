// Layout-identical stand-ins of the reference's fixed-size Eigen float matrices (column-major N x N) used by the
// drop-in headers of the inertial and the mono solvers.
#pragma once

namespace cuvslam {

template <int N>
struct SquareMatrixT {  // column-major N x N
  float m[N * N] = {};
  static SquareMatrixT Zero() { return SquareMatrixT(); }
  static SquareMatrixT Identity() {
    SquareMatrixT r;
    for (int i = 0; i < N; ++i) r.m[i * N + i] = 1.f;
    return r;
  }
  float& operator()(int r, int c) { return m[c * N + r]; }
  float operator()(int r, int c) const { return m[c * N + r]; }
  float* data() { return m; }
  const float* data() const { return m; }
};
using Matrix3T = SquareMatrixT<3>;
using Matrix6T = SquareMatrixT<6>;
using Matrix9T = SquareMatrixT<9>;
using Matrix15T = SquareMatrixT<15>;

}  // namespace cuvslam
