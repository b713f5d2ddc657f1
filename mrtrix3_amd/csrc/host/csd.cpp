#include "csd.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "sh.h"

namespace MR {
namespace DWI {
namespace SDeconv {

namespace {
const double kDirs300[] = {
#include "dirs300.inc"
};
}  // namespace

Matrix electrostatic_repulsion_300() {
  Matrix d(300, 2);
  for (int i = 0; i < 300; ++i) { d(i, 0) = kDirs300[2 * i]; d(i, 1) = kDirs300[2 * i + 1]; }
  return d;
}

Matrix parse_matrix_text(const std::string& text) {
  Matrix out;
  std::vector<std::vector<double>> rows;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    for (char& c : line)
      if (c == ',' || c == ';' || c == '\t' || c == '\r') c = ' ';
    std::istringstream ls(line);
    std::vector<double> r;
    std::string tok;
    while (ls >> tok) {
      char* end = nullptr;
      const double v = std::strtod(tok.c_str(), &end);
      if (end == tok.c_str() || *end) throw std::runtime_error("error converting string \"" + tok + "\" to number");
      r.push_back(v);
    }
    if (r.empty()) continue;
    if (!rows.empty() && r.size() != rows.front().size())
      throw std::runtime_error("uneven rows in matrix");
    rows.push_back(std::move(r));
  }
  if (rows.empty()) return out;
  out = Matrix(int(rows.size()), int(rows.front().size()));
  for (size_t i = 0; i < rows.size(); ++i)
    for (size_t j = 0; j < rows[i].size(); ++j) out(int(i), int(j)) = rows[i][j];
  return out;
}

Matrix load_matrix(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open matrix file \"" + path + "\"");
  std::stringstream ss;
  ss << f.rdbuf();
  Matrix m = parse_matrix_text(ss.str());
  if (m.empty()) throw std::runtime_error("no data in file \"" + path + "\"");
  return m;
}

// math.h:326-336: a vector file is a single row or a single column.
std::vector<double> load_vector(const std::string& path) {
  Matrix m = load_matrix(path);
  if (m.rows != 1 && m.cols != 1)
    throw std::runtime_error("file \"" + path + "\" contains matrix, not vector");
  return m.a;
}

CSD::Shared::Shared(const Matrix& grad_in, const std::vector<double>& requested_shells)
    : grad(grad_in),
      HR_dirs(electrostatic_repulsion_300()),
      neg_lambda(DEFAULT_CSD_NEG_LAMBDA),
      norm_lambda(DEFAULT_CSD_NORM_LAMBDA),
      threshold(DEFAULT_CSD_THRESHOLD),
      lmax_response(0), lmax_data(0), lmax_cmdline(0), lmax(0),
      niter(DEFAULT_CSD_NITER) {
  // b=0 volumes are discarded and exactly one non-zero shell is used (csd.h:57-60)
  dwis = Shells(grad).select_shells(true, false, true, requested_shells).largest().get_volumes();
  DW_dirs = gen_direction_matrix(grad, dwis);
  lmax_data = uint32_t(mrcsd::sh::LforN(int(dwis.size())));
}

void CSD::Shared::parse_cmdline_options(const CSDOptions& opt) {
  if (!opt.lmax.empty()) {
    if (opt.lmax.size() != 1) throw std::runtime_error("CSD algorithm expects a single lmax to be specified");
    lmax_cmdline = opt.lmax.front();
  }
  if (!opt.filter.empty()) init_filter = opt.filter;
  if (!opt.directions.empty()) {
    if (opt.directions.cols != 2) throw std::runtime_error("direction matrix should have 2 columns: [ azimuth elevation ]");
    HR_dirs = opt.directions;
  }
  if (opt.has_neg_lambda) {
    if (opt.neg_lambda < 0.0) throw std::runtime_error("value supplied to -neg_lambda is out of bounds");
    neg_lambda = opt.neg_lambda;
  }
  if (opt.has_norm_lambda) {
    if (opt.norm_lambda < 0.0) throw std::runtime_error("value supplied to -norm_lambda is out of bounds");
    norm_lambda = opt.norm_lambda;
  }
  if (opt.has_threshold) {
    if (opt.threshold < -1.0 || opt.threshold > 10.0) throw std::runtime_error("value supplied to -threshold is out of bounds");
    threshold = opt.threshold;
  }
  if (opt.has_niter) {
    if (opt.niter < 0 || opt.niter > 1000) throw std::runtime_error("value supplied to -niter is out of bounds");
    niter = size_t(opt.niter);
  }
}

void CSD::Shared::set_response(const std::string& path) { set_response(load_vector(path)); }

void CSD::Shared::set_response(const std::vector<double>& coefs) {
  response = coefs;
  lmax_response = response.empty() ? 0 : uint32_t(mrcsd::sh::zonal::LforN(int(response.size())));
}

void CSD::Shared::init() {
  namespace sh = mrcsd::sh;
  if (lmax_data <= 0) throw std::runtime_error("data contain too few directions even for lmax = 2");
  if (lmax_response <= 0) throw std::runtime_error("response function does not contain anisotropic terms");
  lmax = lmax_cmdline ? lmax_cmdline : std::min(lmax_response, DEFAULT_CSD_LMAX);
  if (lmax <= 0 || lmax % 2) throw std::runtime_error("lmax must be a positive even integer");
  lmax_response = std::min(lmax_response, std::min(lmax_data, lmax));

  // linear-deconvolution filter: default [1 1 1], resized to lmax_response/2+1 (csd.h:136-138)
  if (init_filter.empty()) init_filter.assign(3, 1.0);
  init_filter.resize(size_t(sh::zonal::NforL(int(lmax_response))), 0.0);

  std::vector<double> RH = sh::zonal_to_rotational(response);
  if (RH.size() < size_t(sh::zonal::NforL(int(lmax)))) RH.resize(size_t(sh::zonal::NforL(int(lmax))), 0.0);

  // degree index of each SH coefficient: 0 for l=0, 1 for l=2, ...
  auto degree_of = [](int n_coefs) {
    std::vector<int> d(static_cast<size_t>(n_coefs));
    for (int l = 0, c = 0; c < n_coefs; l += 2)
      for (int k = 0; k < 2 * l + 1 && c < n_coefs; ++k) d[size_t(c++)] = l / 2;
    return d;
  };

  // initialisation: regularised inverse of the low-order forward transform (csd.h:145-158)
  Matrix fconv = sh::transform(DW_dirs, int(lmax_response));
  for (int i = 0; i < std::min(fconv.rows, fconv.cols); ++i) fconv(i, i) += 1.0e-2;
  rconv = mrcsd::pinv(fconv);
  {
    const std::vector<int> deg = degree_of(rconv.rows);
    for (int r = 0; r < rconv.rows; ++r) {
      const double f = init_filter[size_t(deg[size_t(r)])] / RH[size_t(deg[size_t(r)])];
      double* row = rconv.row(r);
      for (int c = 0; c < rconv.cols; ++c) row[c] *= f;
    }
  }

  // forward convolution up to lmax (csd.h:163-172)
  fconv = sh::transform(DW_dirs, int(lmax));
  {
    const std::vector<int> deg = degree_of(fconv.cols);
    for (int r = 0; r < fconv.rows; ++r)
      for (int c = 0; c < fconv.cols; ++c) fconv(r, c) *= RH[size_t(deg[size_t(c)])];
  }

  // constraint amplitudes, scaled so that the penalty weight is neg_lambda * 50 * r0 / R
  // (csd.h:175-180)
  HR_trans = sh::transform(HR_dirs, int(lmax));
  const double mult = neg_lambda * 50.0 * response[0] / double(HR_trans.rows);
  for (double& v : HR_trans.a) v *= mult;
  threshold *= mult;

  // normal matrix with the minimum-norm term (csd.h:184-195)
  M = Matrix(DW_dirs.rows, HR_trans.cols, 0.0);
  for (int r = 0; r < M.rows; ++r)
    for (int c = 0; c < fconv.cols; ++c) M(r, c) = fconv(r, c);
  Mt_M = mrcsd::gram(M);
  if (norm_lambda != 0.0) {
    norm_lambda *= NORM_LAMBDA_MULTIPLIER * Mt_M(0, 0);
    for (int i = 0; i < Mt_M.rows; ++i) Mt_M(i, i) += norm_lambda;
  }
}

}  // namespace SDeconv
}  // namespace DWI
}  // namespace MR
