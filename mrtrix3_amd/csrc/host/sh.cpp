#include "sh.h"

#include <stdexcept>

namespace mrcsd {
namespace sh {

namespace {
constexpr double kY00 = 0.282094791773878;   // 1 / (2 sqrt(pi)), legendre.h:110
}

void legendre_column(double* out, int lmax, int m, double x) {
  const double x2 = x * x;
  if (m && x2 >= 1.0) {
    for (int l = m; l <= lmax; ++l) out[l] = 0.0;
    return;
  }
  double seed = kY00;
  if (m) {
    // prod_{k=1..m} (1-x^2)(2k-1)/(2k): the double-factorial ratio of the seed, built up from
    // the innermost factor in the same order as the recursive helper at legendre.h:66-69.
    const double s2 = 1.0 - x2;
    double prod = 1.0;
    for (int k = 1; k <= m; ++k) prod = (s2 * double(2 * k - 1)) / double(2 * k) * prod;
    seed *= std::sqrt(double(2 * m + 1) * prod);
    if (m & 1) seed = -seed;   // Condon-Shortley phase
  }
  out[m] = seed;
  if (lmax == m) return;
  double f = std::sqrt(double(2 * m + 3));
  out[m + 1] = x * f * out[m];
  for (int l = m + 2; l <= lmax; ++l) {
    const double v = x * out[l - 1] - out[l - 2] / f;
    f = std::sqrt(double(4 * l * l - 1) / double(l * l - m * m));
    out[l] = v * f;
  }
}

Matrix transform(const Matrix& dirs, int lmax) {
  if (dirs.cols != 2) throw std::runtime_error("direction matrix should have 2 columns: [ azimuth elevation ]");
  Matrix out(dirs.rows, NforL(lmax));
  std::vector<double> P(size_t(lmax) + 2);
  const double root2 = std::sqrt(2.0);
  for (int d = 0; d < dirs.rows; ++d) {
    const double az = dirs(d, 0);
    const double z = std::cos(dirs(d, 1));
    double* row = out.row(d);
    legendre_column(P.data(), lmax, 0, z);
    for (int l = 0; l <= lmax; l += 2) row[index(l, 0)] = P[l];
    for (int m = 1; m <= lmax; ++m) {
      legendre_column(P.data(), lmax, m, z);
      const double c = std::cos(m * az), s = std::sin(m * az);
      for (int l = (m % 2) ? m + 1 : m; l <= lmax; l += 2) {
        row[index(l, m)] = root2 * P[l] * c;
        row[index(l, -m)] = root2 * P[l] * s;
      }
    }
  }
  return out;
}

std::vector<double> zonal_to_rotational(const std::vector<double>& zsh) {
  const int lmax = zonal::LforN(int(zsh.size()));
  std::vector<double> P(size_t(lmax) + 2);
  legendre_column(P.data(), lmax, 0, 1.0);
  std::vector<double> rh(zsh.size());
  for (int l = 0; l <= lmax; l += 2) rh[size_t(l / 2)] = zsh[size_t(l / 2)] / P[l];
  return rh;
}

}  // namespace sh
}  // namespace mrcsd
