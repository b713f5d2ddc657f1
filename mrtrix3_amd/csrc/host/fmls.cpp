#include "fmls.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <utility>

#include "dense.h"
#include "sh.h"

// The tables are built with the reference's uncontracted arithmetic (no fused multiply-adds): the
// lookup grid's cell corners in particular sit on ties between mirror-image directions, which
// the last bit of a corner's azimuth decides.
#pragma clang fp contract(off)

namespace mrcsd {
namespace fmls {

namespace {

const double kTessellation1281[] = {
#include "dirs1281.inc"
};

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

int slow_nearest(const std::vector<double>& unit, int n, const double* p) {
  int best = 0;
  double mx = std::fabs(dot3(p, &unit[0]));
  for (int i = 1; i < n; ++i) {
    const double d = std::fabs(dot3(p, &unit[size_t(3 * i)]));
    if (d > mx) {
      mx = d;
      best = i;
    }
  }
  return best;
}

size_t grid_index(const DirectionSet& s, const double* p) {
  const double az = std::atan2(p[1], p[0]);
  const double el = std::acos(p[2]);
  const size_t ia = size_t(std::floor((az - s.az_begin) / s.az_step));
  const size_t ie = size_t(std::floor((el - s.el_begin) / s.el_step));
  return ia * size_t(s.n_el) + ie;
}

void build_lookup(DirectionSet& s) {
  // cell size: the mean angle between adjacent directions (set.cpp:372-392)
  double sum = 0.0;
  size_t cnt = 0;
  for (int i = 0; i < s.n; ++i)
    for (int a = 0; a < kMaxAdj; ++a) {
      const int j = s.adj[size_t(i * kMaxAdj + a)];
      if (j > i) {
        sum += std::fabs(dot3(&s.unit[size_t(3 * i)], &s.unit[size_t(3 * j)]));
        ++cnt;
      }
    }
  const double step = std::acos(sum / double(cnt));
  s.n_az = int(std::ceil(2.0 * M_PI / step));
  s.n_el = int(std::ceil(M_PI / step));
  s.az_step = 2.0 * M_PI / double(s.n_az - 1);
  s.el_step = M_PI / double(s.n_el - 1);
  s.az_begin = -M_PI;
  s.el_begin = 0.0;
  const int total = s.n_az * s.n_el;
  std::vector<std::vector<int>> cells(static_cast<size_t>(total));
  for (int i = 0; i < s.n; ++i) cells[grid_index(s, &s.unit[size_t(3 * i)])].push_back(i);
  // the nearest direction to each cell corner (set.cpp:407-429)
  for (int g = 0; g < total; ++g) {
    const int ia = g / s.n_el, ie = g - ia * s.n_el;
    for (int c = 0; c < 4; ++c) {
      double az = s.az_begin + ia * s.az_step, el = s.el_begin + ie * s.el_step;
      if (c == 1 || c == 2) az += s.az_step;
      if (c == 2 || c == 3) el += s.el_step;
      const double p[3] = {std::cos(az) * std::sin(el), std::sin(az) * std::sin(el), std::cos(el)};
      const int nd = slow_nearest(s.unit, s.n, p);
      auto& cell = cells[size_t(g)];
      if (std::find(cell.begin(), cell.end(), nd) == cell.end()) cell.push_back(nd);
    }
  }
  // and the neighbours of every candidate (set.cpp:433-451)
  for (auto& cell : cells) {
    const size_t nexp = cell.size();
    for (size_t e = 0; e < nexp; ++e) {
      const int d = cell[e];
      for (int a = 0; a < kMaxAdj; ++a) {
        const int j = s.adj[size_t(d * kMaxAdj + a)];
        if (j >= 0 && std::find(cell.begin(), cell.end(), j) == cell.end()) cell.push_back(j);
      }
    }
    std::sort(cell.begin(), cell.end());
  }
  s.cell_start.assign(size_t(total) + 1, 0);
  for (int g = 0; g < total; ++g) s.cell_start[size_t(g) + 1] = s.cell_start[size_t(g)] + int(cells[size_t(g)].size());
  s.cell_dirs.clear();
  for (const auto& cell : cells) s.cell_dirs.insert(s.cell_dirs.end(), cell.begin(), cell.end());
}

}  // namespace

std::vector<int> hull_adjacency(const std::vector<double>& unit, int n) {
  // Hull facets of the 2n points {u_i, -u_i}: a triangle (i, a, b) is a facet when no other
  // point lies above its plane.  Facets are small on a well-spread set, so candidates come from
  // the nearest neighbours of each vertex and the test runs over the nearest 48 points; the
  // result is checked against an independent convex hull in the tests.
  // The points are split into ranges over up to 16 host threads, each collecting its own edges;
  // the per-direction neighbour lists are sorted and made unique after, so the result does not
  // depend on the split.
  const int N = 2 * n;
  auto P = [&](int i, int c) { return i < n ? unit[size_t(3 * i + c)] : -unit[size_t(3 * (i - n) + c)]; };
  const int K = std::min(N - 1, 12), Q = std::min(N - 1, 48);
  const int nt = std::max(1, std::min({16, int(std::thread::hardware_concurrency()), N / 128}));
  std::vector<std::vector<std::vector<int>>> nbt(static_cast<size_t>(nt),
                                                 std::vector<std::vector<int>>(static_cast<size_t>(n)));
  auto range = [&](int t) {
    std::vector<std::vector<int>>& nb = nbt[size_t(t)];
    std::vector<std::pair<double, int>> near(static_cast<size_t>(N));
    for (int i = int(int64_t(N) * t / nt); i < int(int64_t(N) * (t + 1) / nt); ++i) {
      const double pi[3] = {P(i, 0), P(i, 1), P(i, 2)};
      for (int j = 0; j < N; ++j) near[size_t(j)] = {j == i ? -2.0 : pi[0] * P(j, 0) + pi[1] * P(j, 1) + pi[2] * P(j, 2), j};
      std::partial_sort(near.begin(), near.begin() + Q, near.end(),
                        [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
                          return x.first > y.first || (x.first == y.first && x.second < y.second);
                        });
      for (int ka = 0; ka < K; ++ka)
        for (int kb = ka + 1; kb < K; ++kb) {
          const int a = near[size_t(ka)].second, b = near[size_t(kb)].second;
          const double ea[3] = {P(a, 0) - pi[0], P(a, 1) - pi[1], P(a, 2) - pi[2]};
          const double eb[3] = {P(b, 0) - pi[0], P(b, 1) - pi[1], P(b, 2) - pi[2]};
          double nr[3] = {ea[1] * eb[2] - ea[2] * eb[1], ea[2] * eb[0] - ea[0] * eb[2], ea[0] * eb[1] - ea[1] * eb[0]};
          const double len = std::sqrt(dot3(nr, nr));
          if (!(len > 0.0)) continue;
          for (double& v : nr) v /= len;
          double d = dot3(nr, pi);
          if (d < 0.0) {
            for (double& v : nr) v = -v;
            d = -d;
          }
          bool facet = true;
          for (int kq = 0; kq < Q && facet; ++kq) {
            const int q = near[size_t(kq)].second;
            if (q == a || q == b) continue;
            if (nr[0] * P(q, 0) + nr[1] * P(q, 1) + nr[2] * P(q, 2) > d + 1e-12) facet = false;
          }
          if (!facet) continue;
          const int v[3] = {i % n, a % n, b % n};
          for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
              if (v[x] != v[y]) nb[size_t(v[x])].push_back(v[y]);
        }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(range, t);
  range(0);
  for (auto& x : th) x.join();
  std::vector<int> adj(size_t(n) * kMaxAdj, -1);
  for (int i = 0; i < n; ++i) {
    std::vector<int> v;
    for (int t = 0; t < nt; ++t) v.insert(v.end(), nbt[size_t(t)][size_t(i)].begin(), nbt[size_t(t)][size_t(i)].end());
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (int(v.size()) > kMaxAdj) throw std::runtime_error("direction set: more than 8 neighbours for a direction");
    std::copy(v.begin(), v.end(), adj.begin() + long(i) * kMaxAdj);
  }
  return adj;
}

std::vector<double> integration_weights(const std::vector<double>& unit, int n) {
  // A w = e: one row per even SH basis function up to lmax, one column per direction, solved in
  // the least-squares sense by Householder QR (the reference uses householderQr, fmls.cpp:141).
  // The reflector of step k is applied to the columns after k by up to 16 host threads, each
  // owning the columns j = t (mod threads) and the right-hand side thread 0; a spin barrier
  // separates the steps.  Every column sees the same operations in the same order as a single
  // thread would apply them, so the weights do not depend on the thread count.
  const int lmax = sh::LforN(n) + 2;
  Matrix dirs(n, 2);
  for (int i = 0; i < n; ++i) {
    dirs(i, 0) = std::atan2(unit[size_t(3 * i + 1)], unit[size_t(3 * i)]);
    dirs(i, 1) = std::acos(unit[size_t(3 * i + 2)]);
  }
  const Matrix S = sh::transform(dirs, lmax);   // [n][m]: column d of A is row d of S
  const int m = S.cols;
  if (m < n) throw std::runtime_error("integration weights: fewer basis functions than directions");
  std::vector<double> A(S.a);                   // column-major A: column d at A[d m .. d m + m)
  std::vector<double> b(size_t(m), 0.0);
  b[0] = 2.0 * std::sqrt(M_PI);
  const int nt = std::max(1, std::min({16, int(std::thread::hardware_concurrency()), n / 64}));
  std::atomic<int> arrived{0}, phase{0};
  std::atomic<bool> failed{false};
  auto barrier = [&]() {
    const int ph = phase.load(std::memory_order_acquire);
    if (arrived.fetch_add(1, std::memory_order_acq_rel) == nt - 1) {
      arrived.store(0, std::memory_order_relaxed);
      phase.store(ph + 1, std::memory_order_release);
    } else {
      while (phase.load(std::memory_order_acquire) == ph) std::this_thread::yield();
    }
  };
  auto worker = [&](int t) {
    std::vector<double> v(static_cast<size_t>(m));
    for (int k = 0; k < n; ++k) {
      // the reflector of column k (final since step k - 1), formed by every thread alike
      const double* ck = &A[size_t(k) * m];
      double nrm2 = 0.0;
      for (int r = k; r < m; ++r) nrm2 += ck[r] * ck[r];
      const double nrm = std::sqrt(nrm2);
      if (nrm == 0.0) failed.store(true);
      const double alpha = ck[k] > 0.0 ? -nrm : nrm;
      for (int r = k; r < m; ++r) v[size_t(r)] = ck[r];
      v[size_t(k)] -= alpha;
      double vv = 0.0;
      for (int r = k; r < m; ++r) vv += v[size_t(r)] * v[size_t(r)];
      const double beta = 2.0 / vv;
      barrier();   // every thread has read column k
      if (k % nt == t) A[size_t(k) * m + k] = alpha;
      if (nrm != 0.0) {
        for (int j = k + 1 + ((t - (k + 1) % nt + nt) % nt); j < n; j += nt) {
          double* cj = &A[size_t(j) * m];
          double s = 0.0;
          for (int r = k; r < m; ++r) s += v[size_t(r)] * cj[r];
          s *= beta;
          for (int r = k; r < m; ++r) cj[r] -= s * v[size_t(r)];
        }
        if (t == 0) {
          double s = 0.0;
          for (int r = k; r < m; ++r) s += v[size_t(r)] * b[size_t(r)];
          s *= beta;
          for (int r = k; r < m; ++r) b[size_t(r)] -= s * v[size_t(r)];
        }
      }
      barrier();   // columns after k and b updated
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(worker, t);
  worker(0);
  for (auto& x : th) x.join();
  if (failed.load()) throw std::runtime_error("integration weights: rank-deficient direction set");
  std::vector<double> w(static_cast<size_t>(n));
  for (int k = n - 1; k >= 0; --k) {
    double s = b[size_t(k)];
    for (int j = k + 1; j < n; ++j) s -= A[size_t(j) * m + k] * w[size_t(j)];
    w[size_t(k)] = s / A[size_t(k) * m + k];
  }
  return w;
}

DirectionSet make_direction_set(const std::vector<double>& az_el) {
  DirectionSet s;
  s.n = int(az_el.size() / 2);
  s.unit.resize(size_t(3 * s.n));
  for (int i = 0; i < s.n; ++i) {
    const double az = az_el[size_t(2 * i)], el = az_el[size_t(2 * i + 1)];
    const double se = std::sin(el);
    s.unit[size_t(3 * i)] = std::cos(az) * se;
    s.unit[size_t(3 * i + 1)] = std::sin(az) * se;
    s.unit[size_t(3 * i + 2)] = std::cos(el);
  }
  s.adj = hull_adjacency(s.unit, s.n);
  s.weights = integration_weights(s.unit, s.n);
  build_lookup(s);
  return s;
}

const DirectionSet& tessellation_1281() {
  static std::once_flag once;
  static DirectionSet set;
  std::call_once(once, [] {
    set = make_direction_set(std::vector<double>(std::begin(kTessellation1281), std::end(kTessellation1281)));
  });
  return set;
}

int select_direction(const DirectionSet& s, const double* p) {
  const size_t g = grid_index(s, p);
  const int b = s.cell_start[g], e = s.cell_start[g + 1];
  int best = s.cell_dirs[size_t(b)];
  double mx = std::fabs(dot3(p, &s.unit[size_t(3 * best)]));
  for (int k = b + 1; k < e; ++k) {
    const int d = s.cell_dirs[size_t(k)];
    const double v = std::fabs(dot3(p, &s.unit[size_t(3 * d)]));
    if (v > mx) {
      mx = v;
      best = d;
    }
  }
  return best;
}

std::vector<double> legendre_table(int lmax, int ndir) {
  const int nAL = (lmax / 2 + 1) * (lmax / 2 + 1);
  std::vector<double> t(size_t(ndir) * nAL, 0.0);
  std::vector<double> buf(size_t(lmax) + 2);
  const double inc = M_PI / (ndir - 1);
  for (int k = 0; k < ndir; ++k) {
    const double x = std::cos(k * inc);
    double* row = &t[size_t(k) * nAL];
    for (int m = 0; m <= lmax; ++m) {
      sh::legendre_column(buf.data(), lmax, m, x);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) row[l * l / 4 + m] = buf[size_t(l)];
    }
  }
  return t;
}

}  // namespace fmls
}  // namespace mrcsd
