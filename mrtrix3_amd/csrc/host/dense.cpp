#include "dense.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace mrcsd {

Matrix Matrix::transpose() const {
  Matrix t(cols, rows);
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) t(c, r) = (*this)(r, c);
  return t;
}

Matrix gram(const Matrix& A) {
  Matrix G(A.cols, A.cols);
  for (int r = 0; r < A.rows; ++r) {
    const double* a = A.row(r);
    for (int i = 0; i < A.cols; ++i) {
      const double ai = a[i];
      double* g = G.row(i);
      for (int j = 0; j <= i; ++j) g[j] += ai * a[j];
    }
  }
  for (int i = 0; i < A.cols; ++i)
    for (int j = 0; j < i; ++j) G(j, i) = G(i, j);
  return G;
}

bool cholesky_lower(Matrix& A) {
  const int n = A.rows;
  for (int j = 0; j < n; ++j) {
    double d = A(j, j);
    for (int k = 0; k < j; ++k) d -= A(j, k) * A(j, k);
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    A(j, j) = d;
    for (int i = j + 1; i < n; ++i) {
      double s = A(i, j);
      for (int k = 0; k < j; ++k) s -= A(i, k) * A(j, k);
      A(i, j) = s / d;
    }
    for (int k = j + 1; k < n; ++k) A(j, k) = 0.0;
  }
  return true;
}

namespace {

// Solve S X = B for symmetric S by LDL^T with diagonal pivoting (largest remaining |d| first).
// S is [n x n], B is [n x m]; returns X.
Matrix ldlt_solve(Matrix S, const Matrix& B) {
  const int n = S.rows;
  std::vector<int> p(static_cast<size_t>(n));
  std::iota(p.begin(), p.end(), 0);
  std::vector<double> col(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(S(i, i)) > std::fabs(S(piv, piv))) piv = i;
    if (piv != k) {
      for (int c = 0; c < n; ++c) std::swap(S(k, c), S(piv, c));
      for (int r = 0; r < n; ++r) std::swap(S(r, k), S(r, piv));
      std::swap(p[size_t(k)], p[size_t(piv)]);
    }
    const double d = S(k, k);
    for (int i = k + 1; i < n; ++i) col[size_t(i)] = S(i, k);
    if (d != 0.0)
      for (int i = k + 1; i < n; ++i) {
        const double f = col[size_t(i)] / d;
        for (int j = k + 1; j < n; ++j) S(i, j) -= f * col[size_t(j)];
      }
    for (int i = k + 1; i < n; ++i) {
      const double l = d != 0.0 ? col[size_t(i)] / d : 0.0;
      S(i, k) = l;
      S(k, i) = l;   // mirrored so later row/column swaps carry L along
    }
  }
  Matrix X(n, B.cols);
  std::vector<double> y(static_cast<size_t>(n));
  for (int c = 0; c < B.cols; ++c) {
    for (int i = 0; i < n; ++i) y[size_t(i)] = B(p[size_t(i)], c);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < i; ++k) y[size_t(i)] -= S(i, k) * y[size_t(k)];
    for (int i = 0; i < n; ++i) y[size_t(i)] = S(i, i) != 0.0 ? y[size_t(i)] / S(i, i) : 0.0;
    for (int i = n - 1; i >= 0; --i)
      for (int k = i + 1; k < n; ++k) y[size_t(i)] -= S(k, i) * y[size_t(k)];
    for (int i = 0; i < n; ++i) X(p[size_t(i)], c) = y[size_t(i)];
  }
  return X;
}

}  // namespace

Matrix pinv(const Matrix& M) {
  if (M.rows >= M.cols) return ldlt_solve(gram(M), M.transpose());
  return ldlt_solve(gram(M.transpose()), M).transpose();
}

Matrix solve_right_lower_transpose(const Matrix& L, const Matrix& Y) {
  // X L^T = Y  <=>  for each row x of X:  sum_k x_k L(j,k) = y_j  (k <= j), forward in j.
  const int n = L.rows;
  Matrix X(Y.rows, n);
  for (int r = 0; r < Y.rows; ++r) {
    double* x = X.row(r);
    const double* y = Y.row(r);
    for (int j = 0; j < n; ++j) {
      double s = y[j];
      for (int k = 0; k < j; ++k) s -= x[k] * L(j, k);
      x[j] = s / L(j, j);
    }
  }
  return X;
}

}  // namespace mrcsd
