#include "icls.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace MR {
namespace Math {
namespace ICLS {

Problem::Problem(const Matrix& problem_matrix, const Matrix& constraint_matrix,
                 const std::vector<double>& constraint_vector, size_t num_equalities,
                 double solution_min_norm_regularisation, double constraint_min_norm_regularisation,
                 size_t max_iterations, double tolerance, bool problem_in_standard_form)
    : H(problem_matrix),
      t(constraint_vector),
      lambda_min_norm(constraint_min_norm_regularisation),
      tol(tolerance),
      max_niter(max_iterations ? max_iterations : size_t(10 * problem_matrix.cols)),
      num_eq(num_equalities),
      standard_form(problem_in_standard_form) {
  if (H.cols != constraint_matrix.cols)
    throw std::runtime_error("FIXME: dimensions of problem and constraint matrices do not match (ICLS)");
  if (solution_min_norm_regularisation < 0.0)
    throw std::runtime_error("FIXME: solution norm regularisation is negative (ICLS)");
  if (lambda_min_norm < 0.0)
    throw std::runtime_error("FIXME: constraint norm regularisation is negative (ICLS)");
  if (tolerance < 0.0) throw std::runtime_error("FIXME: tolerance is negative (ICLS)");
  if (!t.empty() && int(t.size()) != constraint_matrix.rows)
    throw std::runtime_error("FIXME: dimensions of constraint matrix and vector do not match (ICLS)");
  if (num_eq > size_t(constraint_matrix.rows))
    throw std::runtime_error("FIXME: more equality constraints than constraint rows (ICLS)");

  const int n = H.cols;
  // normal matrix (lower triangle), minimum-norm term relative to its largest diagonal entry,
  // Cholesky factor (:123-134)
  chol_HtH = problem_in_standard_form ? H : mrcsd::gram(H);
  double dmax = -INFINITY;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, chol_HtH(i, i));
  for (int i = 0; i < n; ++i) chol_HtH(i, i) += solution_min_norm_regularisation * dmax;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) chol_HtH(i, j) = chol_HtH(j, i);
  if (!mrcsd::cholesky_lower(chol_HtH))
    throw std::runtime_error("ICLS: normal matrix of the problem is not positive definite");

  // b2d = H L^-T (or L^-T in standard form) and B = A L^-T with unit rows, t scaled alike (:139-150)
  if (problem_in_standard_form) {
    Matrix I(n, n, 0.0);
    for (int i = 0; i < n; ++i) I(i, i) = 1.0;
    b2d = mrcsd::solve_right_lower_transpose(chol_HtH, I);
  } else {
    b2d = mrcsd::solve_right_lower_transpose(chol_HtH, H);
  }
  B = mrcsd::solve_right_lower_transpose(chol_HtH, constraint_matrix);
  for (int r = 0; r < B.rows; ++r) {
    double* row = B.row(r);
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += row[j] * row[j];
    const double norm = std::sqrt(s);
    for (int j = 0; j < n; ++j) row[j] /= norm;
    if (!t.empty()) t[size_t(r)] /= norm;
  }
}

}  // namespace ICLS
}  // namespace Math
}  // namespace MR
