#include "gradient.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <stdexcept>

namespace MR {
namespace DWI {

namespace {
double env_or(const char* name, double dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::strtod(v, nullptr);
}
std::string num(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%g", v);
  return buf;
}
}  // namespace

double bzero_threshold() {
  static const double v = env_or("MRTRIX_BZERO_THRESHOLD", 10.0);
  return v;
}
double bvalue_epsilon() {
  static const double v = env_or("MRTRIX_BVALUE_EPSILON", 80.0);
  return v;
}

Matrix get_DW_scheme(const Matrix& raw, int num_volumes, BValueScaling scaling) {
  if (!raw.rows) throw std::runtime_error("no valid diffusion gradient table found");
  if (raw.cols < 4) throw std::runtime_error("unexpected diffusion gradient table matrix dimensions");
  if (num_volumes >= 0 && num_volumes != raw.rows)
    throw std::runtime_error("number of studies in base image (" + std::to_string(num_volumes) +
                             ") does not match number of rows in diffusion gradient table (" +
                             std::to_string(raw.rows) + ")");
  Matrix g = raw;
  std::vector<double> sq(static_cast<size_t>(g.rows));
  double max_log = 0.0;
  for (int r = 0; r < g.rows; ++r) {
    double* v = g.row(r);
    sq[size_t(r)] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (sq[size_t(r)] != 0.0) {
      const double n = std::sqrt(sq[size_t(r)]);
      v[0] /= n; v[1] /= n; v[2] /= n;
    }
    if (sq[size_t(r)] > 0.0) max_log = std::max(max_log, std::fabs(std::log(sq[size_t(r)])));
  }
  const bool needs_scaling = max_log > 0.01;
  if ((needs_scaling && scaling == BValueScaling::Auto) || scaling == BValueScaling::UserOn)
    for (int r = 0; r < g.rows; ++r) g(r, 3) *= sq[size_t(r)];
  return g;
}

Matrix gen_direction_matrix(const Matrix& grad, const std::vector<size_t>& vols) {
  Matrix d(int(vols.size()), 2);
  for (size_t i = 0; i < vols.size(); ++i) {
    const double* g = grad.row(int(vols[i]));
    d(int(i), 0) = std::atan2(g[1], g[0]);
    const double z = g[2] / std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    d(int(i), 1) = z >= 1.0 ? 0.0 : (z <= -1.0 ? M_PI : std::acos(z));
  }
  return d;
}

Shell::Shell(const Matrix& grad, const std::vector<size_t>& volumes)
    : volumes_(volumes), min_(std::numeric_limits<double>::max()) {
  for (size_t v : volumes_) {
    const double b = grad(int(v), 3);
    mean_ += b;
    min_ = std::min(min_, b);
    max_ = std::max(min_, b);   // as written in shells.cpp:64 (tracks min, not a running max)
  }
  mean_ /= double(volumes_.size());
  for (size_t v : volumes_) stdev_ += (grad(int(v), 3) - mean_) * (grad(int(v), 3) - mean_);
  stdev_ = std::sqrt(stdev_ / double(volumes_.size() - 1));
}

// DBSCAN-like clustering of shells.cpp:347-404: b=0 volumes form cluster 1; a non-zero b-value
// seeds a cluster only when at least 3 non-zero b-values lie strictly within epsilon of it.
size_t Shells::cluster_bvalues(const std::vector<double>& b, std::vector<size_t>& clusters) {
  const double b0 = bzero_threshold(), eps = bvalue_epsilon();
  const size_t n = b.size();
  std::vector<char> seen(n, 0);
  size_t id = 0;
  auto neighbours = [&](double ref) {
    std::vector<size_t> out;
    for (size_t k = 0; k < n; ++k)
      if (b[k] > b0 && std::fabs(ref - b[k]) < eps) out.push_back(k);
    return out;
  };
  for (size_t i = 0; i < n; ++i)
    if (b[i] <= b0) { seen[i] = 1; id = 1; clusters[i] = 1; }
  for (size_t i = 0; i < n; ++i) {
    if (seen[i]) continue;
    seen[i] = 1;
    std::vector<size_t> frontier = neighbours(b[i]);
    if (b[i] > b0 && frontier.size() < 3) { clusters[i] = 0; continue; }
    clusters[i] = ++id;
    for (size_t q = 0; q < frontier.size(); ++q) {
      const size_t j = frontier[q];
      if (!seen[j]) {
        seen[j] = 1;
        const std::vector<size_t> more = neighbours(b[j]);
        if (more.size() >= 3) frontier.insert(frontier.end(), more.begin(), more.end());
      }
      if (clusters[j] == 0) clusters[j] = id;
    }
  }
  return id;
}

Shells::Shells(const Matrix& grad) {
  std::vector<double> b(static_cast<size_t>(grad.rows));
  for (int r = 0; r < grad.rows; ++r) b[size_t(r)] = grad(r, 3);
  std::vector<size_t> clusters(b.size(), 0);
  const size_t nclus = cluster_bvalues(b, clusters);
  if (nclus < 1 || double(nclus) > std::sqrt(double(grad.rows)))
    throw std::runtime_error("DWI volumes could not be classified into b-value shells; gradient encoding may not represent a HARDI sequence");
  for (size_t c = 1; c <= nclus; ++c) {
    std::vector<size_t> vols;
    for (size_t v = 0; v < clusters.size(); ++v)
      if (clusters[v] == c) vols.push_back(v);
    shells_.emplace_back(grad, vols);
  }
  std::sort(shells_.begin(), shells_.end());
}

Shells& Shells::select_shells(bool force_singleshell, bool force_with_bzero, bool force_without_bzero,
                              const std::vector<double>& requested) {
  if (force_without_bzero && force_with_bzero)
    throw std::runtime_error("Incompatible constraints: command tries to enforce proceeding both with and without b=0");
  std::vector<char> keep(count(), 0);
  if (!requested.empty()) {
    bool bzero_selected = false;
    size_t nonzero_selected = 0;
    for (double b : requested) {
      if (b < 0) throw std::runtime_error("Cannot select shells corresponding to negative b-values");
      if (b <= bzero_threshold()) {
        if (!has_bzero()) throw std::runtime_error("User selected b=0 shell, but no such data was found");
        if (bzero_selected) throw std::runtime_error("User selected b=0 shell more than once");
        keep[0] = 1;
        bzero_selected = true;
        continue;
      }
      bool selected = false;
      for (size_t s = 0; s < count(); ++s)
        if (b >= shells_[s].get_min() && b <= shells_[s].get_max()) {
          if (keep[s]) throw std::runtime_error("User selected a shell more than once");
          keep[s] = 1; ++nonzero_selected; selected = true;
        }
      if (selected) continue;
      size_t best = 0; bool ambiguous = false;
      for (size_t s = 0; s < count(); ++s)
        if (std::fabs(b - shells_[s].get_mean()) <= 1.0) {
          if (selected) ambiguous = true; else { best = s; selected = true; }
        }
      if (selected && !ambiguous) {
        if (keep[best]) throw std::runtime_error("User selected a shell more than once");
        keep[best] = 1; ++nonzero_selected;
        continue;
      }
      bool zero_stdev = false;
      for (const auto& s : shells_)
        if (!s.is_bzero() && s.get_stdev() < 1.0) { zero_stdev = true; break; }
      double best_n = std::numeric_limits<double>::max();
      best = 0; ambiguous = false;
      for (size_t s = 0; s < count(); ++s) {
        const double sd = shells_[s].is_bzero() ? 0.5 * bzero_threshold()
                          : (zero_stdev ? std::sqrt(shells_[s].get_mean()) : shells_[s].get_stdev());
        const double ns = std::fabs((b - shells_[s].get_mean()) / sd);
        if (ns < best_n) { ambiguous = ns >= 0.1 * best_n; best = s; best_n = ns; }
        else ambiguous = ns < 10.0 * best_n;
      }
      if (ambiguous) throw std::runtime_error("Unable to robustly select desired shell b=" + num(b));
      if (keep[best]) throw std::runtime_error("User selected a shell more than once");
      keep[best] = 1; ++nonzero_selected;
    }
    if (force_singleshell && nonzero_selected != 1)
      throw std::runtime_error("User selected " + std::to_string(nonzero_selected) +
                               " non b=0 shells, but the command requires single-shell data");
    if (force_with_bzero && !bzero_selected)
      throw std::runtime_error("User did not select b=0 shell, but the command requires the presence of b=0 data");
    if (force_without_bzero && bzero_selected)
      throw std::runtime_error("User selected b=0 shell, but the command is not compatible with b=0 data");
  } else {
    if (force_singleshell && !is_single_shell()) {
      if (count() == 1 && has_bzero())
        throw std::runtime_error("No non b=0 data found, but the command requires a non b=0 shell");
      keep[count() - 1] = 1;
      if (has_bzero()) keep[0] = 1;
    } else {
      std::fill(keep.begin(), keep.end(), 1);
    }
    if (force_with_bzero && !has_bzero())
      throw std::runtime_error("No b=0 data found, but the command requires the presence of b=0 data");
    if (force_without_bzero && has_bzero()) keep[0] = 0;
  }
  std::vector<Shell> kept;
  for (size_t s = 0; s < count(); ++s)
    if (keep[s]) kept.push_back(shells_[s]);
  shells_.swap(kept);
  if (shells_.empty()) throw std::runtime_error("No DW shells remain after shell selection");
  return *this;
}

}  // namespace DWI
}  // namespace MR
