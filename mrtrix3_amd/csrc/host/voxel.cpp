// Per-voxel solver classes with the reference's call surface, backed by the GPU kernels:
//   MR::DWI::SDeconv::CSD        (src/dwi/sdeconv/csd.h:214-280)
//   MR::Math::ICLS::Solver       (core/math/constrained_least_squares.h:225-365)
// (MSMT_CSD, msmt_csd.h:262-275, is inline in msmt.h on top of ICLS::Solver.)
// Each object solves one voxel per launch on its own HIP stream, so objects used from several
// host threads (one copy per thread, as ThreadedLoop makes them) run concurrently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../capi_csd_ctx.h"
#include "../../../include/mricls.h"
#include "../capi_icls_ctx.h"
#include "csd.h"
#include "icls.h"

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

hipStream_t make_stream() {
  hipStream_t s = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  return s;
}

}  // namespace

namespace MR {
namespace DWI {
namespace SDeconv {

// The problem matrices on the device, with the voxel's signals as a row of exactly the selected
// volumes (volume index i for signal i).
struct CSD::Engine {
  mrcsd_ctx* ctx = nullptr;
  int n_dw = 0, n_sh = 0;
  ~Engine() { mrcsd_destroy(ctx); }
};

struct CSD::Voxel {
  int device = 0;
  hipStream_t stream = nullptr;
  mrcsd::DeviceBuffer<float> dw, fod;
  mrcsd::DeviceBuffer<uint16_t> iters;
  mrcsd::DeviceBuffer<unsigned long long> counter;
  ~Voxel() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamDestroy(stream);
    }
  }
};

CSD::CSD(const Shared& shared_data) : shared(shared_data) {
  if (shared.nSH() == 0 || shared.dwis.empty()) throw std::runtime_error("CSD: shared data not initialised (call init())");
  auto e = std::make_shared<Engine>();
  std::vector<int> dwis(shared.dwis.size());
  for (size_t i = 0; i < dwis.size(); ++i) dwis[i] = int(i);
  mrcsd_params p;
  p.n_dw = int(shared.dwis.size());
  p.n_sh = int(shared.nSH());
  p.n_init = shared.rconv.rows;
  p.n_hr = shared.HR_trans.rows;
  p.rconv = shared.rconv.data();
  p.M = shared.M.data();
  p.MtM = shared.Mt_M.data();
  p.HR_trans = shared.HR_trans.data();
  p.dwis = dwis.data();
  p.threshold = shared.threshold;
  p.niter = int(std::min<size_t>(shared.niter, 65534));
  if (mrcsd_create(&p, &e->ctx) != MRCSD_OK) throw std::runtime_error(std::string("CSD: ") + mrcsd_last_error());
  e->n_dw = p.n_dw;
  e->n_sh = p.n_sh;
  engine_ = std::move(e);
  F_.assign(shared.nSH(), 0.0);
}

CSD::CSD(const CSD& other)
    : shared(other.shared),
      engine_(other.engine_),
      b_(other.b_),
      calls_(other.calls_),
      F_(other.F_),
      solved_budget_(other.solved_budget_),
      solved_n_(other.solved_n_),
      F_state_(other.F_state_) {}

CSD::~CSD() = default;

void CSD::set(const std::vector<double>& DW_signals) {
  if (DW_signals.size() != size_t(engine_->n_dw))
    throw std::runtime_error("CSD::set: expected " + std::to_string(engine_->n_dw) + " DW signals, got " +
                             std::to_string(DW_signals.size()));
  b_.resize(DW_signals.size());
  for (size_t i = 0; i < b_.size(); ++i) b_[i] = float(DW_signals[i]);
  calls_ = 0;
  solved_budget_ = 0;
  solved_n_ = 0;
  F_state_ = size_t(-1);
}

// One GPU launch for this voxel with an iteration budget: F after min(budget, n) solves into F_,
// and the loop counter n at which the negative set stopped changing (n == budget: not within it).
// Returns n, or SIZE_MAX for a voxel with a non-finite signal (zero FOD, as dwi2fod writes it).
size_t CSD::solve(size_t budget) const {
  if (b_.empty()) throw std::runtime_error("CSD: set() must be called before iterate() or FOD()");
  const mrcsd_ctx* ctx = engine_->ctx;
  if (!dev_) {
    auto v = std::make_unique<Voxel>();
    v->device = ctx->device;
    hip_check(hipSetDevice(v->device), "hipSetDevice");
    v->stream = make_stream();
    v->dw.alloc(size_t(engine_->n_dw));
    v->fod.alloc(size_t(engine_->n_sh));
    v->iters.alloc(1);
    v->counter.alloc(1);
    dev_ = std::move(v);
  }
  Voxel& v = *dev_;
  hip_check(hipSetDevice(v.device), "hipSetDevice");
  hip_check(hipMemcpyAsync(v.dw.get(), b_.data(), b_.size() * sizeof(float), hipMemcpyHostToDevice, v.stream),
            "hipMemcpyAsync");
  mrcsd::CsdDeviceProblem P = ctx->P;
  P.niter = int(budget);
  mrcsd::CsdLaunch L{};
  L.dw = v.dw.get();
  L.ld_dw = engine_->n_dw;
  L.nvox = 1;
  L.fod = v.fod.get();
  L.iters = v.iters.get();
  L.counter = v.counter.get();
  hip_check(mrcsd::csd_launch(P, L, v.stream), "csd_launch");
  std::vector<float> f(size_t(engine_->n_sh));
  uint16_t it = 0;
  hip_check(hipMemcpyAsync(f.data(), v.fod.get(), f.size() * sizeof(float), hipMemcpyDeviceToHost, v.stream),
            "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(&it, v.iters.get(), sizeof(uint16_t), hipMemcpyDeviceToHost, v.stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(v.stream), "hipStreamSynchronize");
  for (size_t i = 0; i < f.size(); ++i) F_[i] = double(f[i]);
  return it == 0xFFFF ? size_t(-1) : size_t(it);
}

bool CSD::iterate() {
  const size_t c = calls_++;   // this is iterate() call c since set()
  const bool known = solved_budget_ > 0 && (solved_n_ < solved_budget_ || solved_budget_ > c);
  if (!known) {
    const size_t budget = std::min<size_t>(std::max<size_t>(shared.niter, 2 * (c + 1)), 65534);
    const size_t n = solve(budget);
    if (n == size_t(-1)) {    // non-finite signal: nothing to iterate
      solved_budget_ = 1;
      solved_n_ = 0;
    } else {
      solved_budget_ = budget;
      solved_n_ = n;
    }
    F_state_ = solved_n_ < solved_budget_ ? solved_n_ : solved_budget_;
  }
  return solved_n_ < solved_budget_ && solved_n_ <= c;
}

const std::vector<double>& CSD::FOD() const {
  // after calls_ iterate() calls F is the reference's F after min(calls_, n) solves
  const bool converged = solved_budget_ > 0 && solved_n_ < solved_budget_;
  const size_t want = converged ? std::min(calls_, solved_n_) : calls_;
  if (F_state_ != want) {
    solve(std::min<size_t>(want, 65534));   // F mid-way: the same trajectory stopped after `want` solves
    F_state_ = want;
  }
  return F_;
}

}  // namespace SDeconv
}  // namespace DWI

namespace Math {
namespace ICLS {

struct Solver::Engine {
  mricls_ctx* ctx = nullptr;
  ~Engine() { mricls_destroy(ctx); }
};

struct Solver::Voxel {
  int device = 0;
  hipStream_t stream = nullptr;
  mrcsd::DeviceBuffer<double> b, x;
  mrcsd::DeviceBuffer<uint16_t> out;   // iters, nactive
  mrcsd::DeviceBuffer<unsigned long long> counters;
  mrcsd::DeviceBuffer<uint32_t> queue;
  ~Voxel() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamDestroy(stream);
    }
  }
};

Solver::Solver(const Problem& problem) : P(problem) {
  auto e = std::make_shared<Engine>();
  if (mrcsd::icls_ctx_from_problem(P, &e->ctx) != MRCSD_OK)
    throw std::runtime_error(std::string("ICLS::Solver: ") + mrcsd_last_error());
  engine_ = std::move(e);
}

Solver::Solver(const Solver& other) : P(other.P), engine_(other.engine_), nactive_(other.nactive_) {}

Solver::~Solver() = default;

size_t Solver::operator()(std::vector<double>& x, const std::vector<double>& b) {
  mricls_ctx* ctx = engine_->ctx;
  const int nm = ctx->P.nmeas, n = ctx->P.n;
  if (b.size() != size_t(nm))
    throw std::runtime_error("ICLS::Solver: expected " + std::to_string(nm) + " measurements, got " +
                             std::to_string(b.size()));
  if (!dev_) {
    auto v = std::make_unique<Voxel>();
    v->device = ctx->device;
    hip_check(hipSetDevice(v->device), "hipSetDevice");
    v->stream = make_stream();
    v->b.alloc(size_t(nm));
    v->x.alloc(size_t(n));
    v->out.alloc(2);
    v->counters.alloc(2 * mrcsd::kIclsMaxPasses);
    v->queue.alloc(2);
    dev_ = std::move(v);
  }
  Voxel& v = *dev_;
  hip_check(hipSetDevice(v.device), "hipSetDevice");
  hip_check(hipMemcpyAsync(v.b.get(), b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice, v.stream),
            "hipMemcpyAsync");
  mrcsd::IclsScratch S{};
  S.counters = v.counters.get();
  S.queue[0] = v.queue.get();
  S.queue[1] = v.queue.get() + 1;
  S.capacity = 1;
  if (mrcsd::icls_solve_impl(ctx, nullptr, v.b.get(), nm, nullptr, 1, nullptr, v.x.get(), v.out.get(), v.out.get() + 1,
                             nullptr, &S, v.stream, "ICLS::Solver") != MRCSD_OK)
    throw std::runtime_error(mrcsd_last_error());
  x.resize(size_t(n));
  uint16_t out[2] = {0, 0};
  hip_check(hipMemcpyAsync(x.data(), v.x.get(), size_t(n) * sizeof(double), hipMemcpyDeviceToHost, v.stream),
            "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(out, v.out.get(), sizeof(out), hipMemcpyDeviceToHost, v.stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(v.stream), "hipStreamSynchronize");
  if (out[0] == 0xFFFE)
    throw std::runtime_error("ICLS::Solver: more than 128 active constraints (not supported by the GPU solver)");
  nactive_ = out[1];
  return out[0];
}

}  // namespace ICLS
}  // namespace Math
}  // namespace MR
