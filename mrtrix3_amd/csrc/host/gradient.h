// Diffusion gradient table handling and b-value shell clustering (host side, set-up only).
// Restates the behaviour of core/dwi/gradient.{h,cpp} and core/dwi/shells.{h,cpp}.
#pragma once

#include <string>
#include <vector>

#include "dense.h"

namespace MR {
namespace DWI {

using mrcsd::Matrix;

// gradient.h:30 and the BZeroThreshold / BValueEpsilon config keys (gradient.cpp:88-95,
// shells.cpp:43-46); set from the environment variables MRTRIX_BZERO_THRESHOLD and
// MRTRIX_BVALUE_EPSILON when present.
double bzero_threshold();
double bvalue_epsilon();

enum class BValueScaling { Auto, UserOn, UserOff };

// get_DW_scheme (gradient.cpp:280-341) applied to an already-parsed [nvol x 4] table:
// validates it against the number of volumes (gradient.h:52-67), normalises the direction
// rows and scales b by |g|^2 when any |log |g|^2| exceeds 0.01 (or when forced).
Matrix get_DW_scheme(const Matrix& raw, int num_volumes, BValueScaling scaling = BValueScaling::Auto);

// gradient.h:77-94: [azimuth elevation] of the listed rows.
Matrix gen_direction_matrix(const Matrix& grad, const std::vector<size_t>& vols);

class Shell {
 public:
  Shell() = default;
  Shell(const Matrix& grad, const std::vector<size_t>& volumes);   // shells.cpp:52-70
  const std::vector<size_t>& get_volumes() const { return volumes_; }
  size_t count() const { return volumes_.size(); }
  double get_mean() const { return mean_; }
  double get_stdev() const { return stdev_; }
  double get_min() const { return min_; }
  double get_max() const { return max_; }
  bool is_bzero() const { return mean_ <= bzero_threshold(); }
  bool operator<(const Shell& rhs) const { return mean_ < rhs.mean_; }

 private:
  std::vector<size_t> volumes_;
  double mean_ = 0.0, stdev_ = 0.0, min_ = 0.0, max_ = 0.0;
};

class Shells {
 public:
  explicit Shells(const Matrix& grad);                               // shells.cpp:300-342
  const Shell& operator[](size_t i) const { return shells_[i]; }
  const Shell& smallest() const { return shells_.front(); }
  const Shell& largest() const { return shells_.back(); }
  size_t count() const { return shells_.size(); }
  bool has_bzero() const { return smallest().is_bzero(); }
  bool is_single_shell() const {
    return (count() == 1 && !has_bzero()) || (count() == 2 && has_bzero());
  }
  // shells.cpp:100-281.  `requested` holds the b-values of a -shells option (empty: none).
  Shells& select_shells(bool force_singleshell, bool force_with_bzero, bool force_without_bzero,
                        const std::vector<double>& requested = {});

 private:
  std::vector<Shell> shells_;
  static size_t cluster_bvalues(const std::vector<double>& b, std::vector<size_t>& clusters);
};

}  // namespace DWI
}  // namespace MR
