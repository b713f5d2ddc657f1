// Real even-order spherical harmonics used to set up the CSD / MSMT-CSD problems on the host.
// Conventions follow MRtrix3 (docs/concepts/spherical_harmonics.rst:74-134): orthonormal basis,
// Condon-Shortley phase in the associated Legendre functions, sqrt(2) on m != 0 terms, cosine
// terms at +m and sine terms at -m, index(l,m) = l(l+1)/2 + m over even l only.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

#include "dense.h"

namespace mrcsd {
namespace sh {

// core/math/SH.h:46-49
inline int NforL(int lmax) { return (lmax + 1) * (lmax + 2) / 2; }
// core/math/SH.h:52-55
inline int index(int l, int m) { return l * (l + 1) / 2 + m; }
// core/math/SH.h:70-73 (the reference evaluates the square root in single precision)
inline int LforN(int N) {
  return N ? 2 * int(std::floor((std::sqrt(float(1 + 8 * N)) - 3.0) / 4.0)) : 0;
}

namespace zonal {
// core/math/ZSH.h:42-57
inline int NforL(int lmax) { return 1 + lmax / 2; }
inline int LforN(int N) { return 2 * (N - 1); }
}  // namespace zonal

// Orthonormal associated Legendre functions P_l^m(x) for fixed m and l = m..lmax, stored at
// out[l].  Same seed, sign convention and three-term recursion as core/math/legendre.h:100-123.
void legendre_column(double* out, int lmax, int m, double x);

// Amplitude matrix [ndir x NforL(lmax)] for directions given as [azimuth elevation] rows
// (core/math/SH.h:81-104).
Matrix transform(const Matrix& dirs_az_el, int lmax);

// Zonal response coefficients to rotational harmonics: rh_l = r_l / P_l^0(1)
// (core/math/ZSH.h:220-231).
std::vector<double> zonal_to_rotational(const std::vector<double>& zsh);

}  // namespace sh
}  // namespace mrcsd
