// Host-side set-up of the inequality-constrained least-squares problem
// (Math::ICLS::Problem, core/math/constrained_least_squares.h:73-197).  The per-voxel solver
// (Solver::operator(), :237-357) runs on the GPU, see kernels/icls_solve.hip.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "dense.h"

namespace MR {
namespace Math {
namespace ICLS {

using mrcsd::Matrix;

class Problem {
 public:
  Problem() = default;
  // :91-152.  Equality rows (if any) are the last num_equalities rows of the constraint matrix;
  // an empty t means zeros.  max_iterations 0 means 10 * ncols.
  Problem(const Matrix& problem_matrix, const Matrix& constraint_matrix,
          const std::vector<double>& constraint_vector = {}, size_t num_equalities = 0,
          double solution_min_norm_regularisation = 0.0,
          double constraint_min_norm_regularisation = 0.0, size_t max_iterations = 0,
          double tolerance = 0.0, bool problem_in_standard_form = false);

  size_t num_parameters() const { return size_t(H.cols); }
  size_t num_measurements() const { return size_t(b2d.rows); }
  size_t num_constraints() const { return size_t(B.rows); }
  size_t num_equalities() const { return num_eq; }

  Matrix H, chol_HtH, B, b2d;
  std::vector<double> t;
  double lambda_min_norm = 0.0, tol = 0.0;
  size_t max_niter = 0, num_eq = 0;
  bool standard_form = false;   // H holds H^T H (the problem was given in standard form)
};

// Solver::operator()(x, b) (:225-359) for one right-hand side, run on the GPU in double
// precision (one voxel per launch; batches belong to mricls_solve in include/mricls.h).  Returns
// the reference's niter and leaves the final active set size in active_constraints().  The
// device copy of the problem is made at construction and shared by copies; each object owns its
// device buffers and HIP stream.  A right-hand side whose active set would exceed 128
// constraints throws (the reference has no such limit).
class Solver {
 public:
  explicit Solver(const Problem& problem);
  Solver(const Solver& other);
  Solver& operator=(const Solver&) = delete;
  ~Solver();

  size_t operator()(std::vector<double>& x, const std::vector<double>& b);
  const Problem& problem() const { return P; }
  size_t active_constraints() const { return nactive_; }

  struct Engine;
  struct Voxel;

 private:
  const Problem& P;
  std::shared_ptr<Engine> engine_;
  std::unique_ptr<Voxel> dev_;
  size_t nactive_ = 0;
};

}  // namespace ICLS
}  // namespace Math
}  // namespace MR
