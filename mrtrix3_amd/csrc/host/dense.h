// Minimal dense FP64 matrices for one-off problem set-up on the host (a few 100x100 at most).
// Row-major storage; nothing here is on the per-voxel path, which runs on the GPU.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace mrcsd {

struct Matrix {
  int rows = 0, cols = 0;
  std::vector<double> a;

  Matrix() = default;
  Matrix(int r, int c, double v = 0.0) : rows(r), cols(c), a(size_t(r) * size_t(c), v) {}

  double& operator()(int r, int c) { return a[size_t(r) * size_t(cols) + size_t(c)]; }
  double operator()(int r, int c) const { return a[size_t(r) * size_t(cols) + size_t(c)]; }
  double* row(int r) { return a.data() + size_t(r) * size_t(cols); }
  const double* row(int r) const { return a.data() + size_t(r) * size_t(cols); }
  double* data() { return a.data(); }
  const double* data() const { return a.data(); }
  bool empty() const { return a.empty(); }

  Matrix transpose() const;
};

// C = A^T A (full symmetric result).
Matrix gram(const Matrix& A);

// In-place Cholesky of the lower triangle of a symmetric matrix; upper triangle zeroed.
// Returns false on a non-positive pivot.
bool cholesky_lower(Matrix& A);

// Pseudo-inverse by the normal equations, solved with a symmetric-pivoted LDL^T:
// tall: (M^T M)^-1 M^T, wide: ((M M^T)^-1 M)^T.  Same construction as core/math/least_squares.h:38-45.
Matrix pinv(const Matrix& M);

// X = Y L^-T for lower-triangular L (the "solve on the right" of constrained_least_squares.h:139-144).
Matrix solve_right_lower_transpose(const Matrix& L, const Matrix& Y);

}  // namespace mrcsd
