#include "msmt.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "csd.h"
#include "sh.h"

namespace MR {
namespace DWI {
namespace SDeconv {

MSMT_CSD::Shared::Shared(const Matrix& g, const std::vector<double>& requested_shells)
    : grad(g),
      shells(g),
      HR_dirs(electrostatic_repulsion_300()),
      solution_min_norm_regularisation(DEFAULT_MSMTCSD_NORM_LAMBDA),
      constraint_min_norm_regularisation(DEFAULT_MSMTCSD_NEG_LAMBDA) {
  shells.select_shells(false, false, false, requested_shells);
}

void MSMT_CSD::Shared::set_responses(const std::vector<Matrix>& matrices) {
  responses = matrices;
  prepare_responses();
}

void MSMT_CSD::Shared::set_responses(const std::vector<std::string>& files) {
  std::vector<Matrix> m;
  for (const auto& f : files) {
    try {
      m.push_back(load_matrix(f));
    } catch (const std::exception& e) {
      throw std::runtime_error("File \"" + f + "\" is not a valid response function file: " + e.what());
    }
  }
  response_files = files;
  set_responses(m);
}

void MSMT_CSD::Shared::prepare_responses() {
  lmax_response.clear();
  for (Matrix& r : responses) {
    // drop trailing degrees whose coefficients are zero in every shell
    int n = 0;
    for (int row = 0; row < r.rows; ++row)
      for (int col = 0; col < r.cols; ++col)
        if (r(row, col) != 0.0) n = std::max(n, col + 1);
    Matrix t(r.rows, n);
    for (int row = 0; row < r.rows; ++row)
      for (int col = 0; col < n; ++col) t(row, col) = r(row, col);
    r = t;
    lmax_response.push_back(uint32_t(mrcsd::sh::zonal::LforN(n)));
  }
}

void MSMT_CSD::Shared::init() {
  namespace sh = mrcsd::sh;
  const size_t nt = num_tissues();
  if (lmax.empty()) {
    lmax = lmax_response;
    for (auto& l : lmax) l = std::min(DEFAULT_MSMTCSD_LMAX, l);
  } else {
    if (lmax.size() != nt)
      throw std::runtime_error("Number of lmaxes specified (" + std::to_string(lmax.size()) +
                               ") does not match number of tissues (" + std::to_string(nt) + ")");
    for (auto l : lmax)
      if (l % 2) throw std::runtime_error("Each value of lmax must be a non-negative even integer");
  }
  for (size_t t = 0; t < nt; ++t) {
    if (size_t(responses[t].rows) != num_shells())
      throw std::runtime_error("number of rows in response functions must match number of b-value shells; "
                               "number of shells is " + std::to_string(num_shells()) + ", but file \"" +
                               (t < response_files.size() ? response_files[t] : std::string("(matrix)")) +
                               "\" contains " + std::to_string(responses[t].rows) + " rows");
    // pad (or cut) to the requested lmax of this tissue
    const int want = sh::zonal::NforL(int(lmax[t]));
    Matrix p(responses[t].rows, want, 0.0);
    for (int r = 0; r < p.rows; ++r)
      for (int c = 0; c < std::min(want, responses[t].cols); ++c) p(r, c) = responses[t](r, c);
    responses[t] = p;
  }

  int nparams = 0;
  uint32_t maxlmax = 0;
  for (size_t t = 0; t < nt; ++t) {
    nparams += sh::NforL(int(lmax[t]));
    maxlmax = std::max(maxlmax, lmax[t]);
  }

  // SH of every volume direction; b=0 rows have an undefined elevation and keep only Y_00
  std::vector<size_t> all(size_t(grad.rows));
  for (size_t i = 0; i < all.size(); ++i) all[i] = i;
  Matrix SHT = sh::transform(gen_direction_matrix(grad, all), int(maxlmax));
  for (double& v : SHT.a)
    if (std::isnan(v)) v = 0.0;
  // P_l^0(1) per even degree (the non-zero entries of the SH row at the pole)
  std::vector<double> dsh;
  {
    Matrix pole(1, 2, 0.0);
    Matrix row = sh::transform(pole, int(maxlmax));
    for (double v : row.a)
      if (v != 0.0) dsh.push_back(v);
  }

  C = Matrix(grad.rows, nparams, 0.0);
  int pbegin = 0;
  for (size_t t = 0; t < nt; ++t) {
    const int tl = int(lmax[t]), tn = sh::NforL(tl);
    for (size_t s = 0; s < num_shells(); ++s) {
      std::vector<double> conv(static_cast<size_t>(tn));
      for (int l = 0, idx = 0; l <= tl; l += 2)
        for (int m = -l; m <= l; ++m) conv[size_t(idx++)] = responses[t](int(s), l / 2) / dsh[size_t(l / 2)];
      for (size_t v : shells[s].get_volumes())
        for (int k = 0; k < tn; ++k) C(int(v), pbegin + k) = SHT(int(v), k) * conv[size_t(k)];
    }
    pbegin += tn;
  }

  Matrix HR = sh::transform(HR_dirs, int(maxlmax));
  int M = 0;
  for (size_t t = 0; t < nt; ++t) M += lmax[t] > 0 ? HR_dirs.rows : 1;
  A = Matrix(M, nparams, 0.0);
  int bm = 0, bn = 0;
  for (size_t t = 0; t < nt; ++t) {
    const int mt = lmax[t] > 0 ? HR_dirs.rows : 1, ntc = sh::NforL(int(lmax[t]));
    for (int i = 0; i < mt; ++i)
      for (int j = 0; j < ntc; ++j) A(bm + i, bn + j) = HR(i, j);
    bm += mt;
    bn += ntc;
  }
  problem = Math::ICLS::Problem(C, A, {}, 0, solution_min_norm_regularisation,
                                constraint_min_norm_regularisation);
}

}  // namespace SDeconv
}  // namespace DWI
}  // namespace MR
