#include "gt.h"

#include <cmath>
#include <stdexcept>
#include <string>

#include "gradient.h"
#include "sh.h"

namespace MR {
namespace DWI {
namespace Tractography {
namespace GT {

std::vector<double> sh_delta(const double* u, int lmax) {
  using mrcsd::sh::index;
  std::vector<double> out(size_t(mrcsd::sh::NforL(lmax)), 0.0), AL(size_t(lmax) + 2);
  const double rxy = std::sqrt(u[1] * u[1] + u[0] * u[0]);
  const double cp = rxy != 0.0 ? u[0] / rxy : 1.0;
  const double sp = rxy != 0.0 ? u[1] / rxy : 0.0;
  const double root2 = std::sqrt(2.0);
  mrcsd::sh::legendre_column(AL.data(), lmax, 0, u[2]);
  for (int l = 0; l <= lmax; l += 2) out[size_t(index(l, 0))] = AL[size_t(l)];
  // cos(m az), sin(m az) by the angle-addition recurrence from (cp, sp)
  double c0 = 1.0, s0 = 0.0;
  for (int m = 1; m <= lmax; ++m) {
    mrcsd::sh::legendre_column(AL.data(), lmax, m, u[2]);
    const double c = c0 * cp - s0 * sp;
    const double s = s0 * cp + c0 * sp;
    for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) {
      out[size_t(index(l, m))] = AL[size_t(l)] * root2 * c;
      out[size_t(index(l, -m))] = AL[size_t(l)] * root2 * s;
    }
    c0 = c;
    s0 = s;
  }
  return out;
}

void external_energy_kernels(const Matrix& grad, const Matrix& resp_wm, const std::vector<std::vector<double>>& resp_iso,
                             int lmax, double weight, Matrix& K, Matrix& Ak) {
  if (lmax < 0 || (lmax & 1)) throw std::runtime_error("lmax must be a non-negative even integer");
  const int nrows = grad.rows, ncols = mrcsd::sh::NforL(lmax), nf = int(resp_iso.size());
  const Shells shells(grad);
  if (size_t(resp_wm.rows) != shells.count())
    throw std::runtime_error("WM kernel size does not match the no. b-values in the image.");
  for (const auto& r : resp_iso)
    if (r.size() != shells.count())
      throw std::runtime_error("Isotropic kernel size does not match the no. b-values in the image.");
  K = Matrix(nrows, ncols, 0.0);
  Ak = Matrix(nrows, nf + 1, 0.0);
  const int nz = mrcsd::sh::zonal::NforL(lmax);
  const double sqrt4pi = std::sqrt(4.0 * M_PI);
  for (size_t s = 0; s < shells.count(); ++s) {
    std::vector<double> zsh(size_t(nz), 0.0);
    for (int i = 0; i < nz; ++i) zsh[size_t(i)] = i < resp_wm.cols ? resp_wm(int(s), i) : 0.0;
    const std::vector<double> rh = mrcsd::sh::zonal_to_rotational(zsh);
    const double wmr0 = resp_wm(int(s), 0) / sqrt4pi;
    for (size_t r : shells[s].get_volumes()) {
      double u[3] = {grad(int(r), 0), grad(int(r), 1), grad(int(r), 2)};
      const double n = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      if (n > 0.0)
        for (double& v : u) v /= n;
      std::vector<double> d = sh_delta(u, lmax);
      // sconv (SH.h:349-359): every coefficient of degree l times rh_l
      for (int l = 0; l <= lmax; l += 2)
        for (int m = -l; m <= l; ++m) d[size_t(mrcsd::sh::index(l, m))] *= rh[size_t(l / 2)];
      for (int c = 0; c < ncols; ++c) K(int(r), c) = d[size_t(c)];
      Ak(int(r), 0) = wmr0;
      for (int j = 0; j < nf; ++j) Ak(int(r), j + 1) = resp_iso[size_t(j)][s];
    }
  }
  for (double& v : K.a) v *= weight;
}

}  // namespace GT
}  // namespace Tractography
}  // namespace DWI
}  // namespace MR
