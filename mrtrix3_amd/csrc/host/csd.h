// Host side of the single-tissue constrained spherical deconvolution path.
//
// MR::DWI::SDeconv::CSD::Shared keeps the call surface of the reference class
// (src/dwi/sdeconv/csd.h:49-212): construct from a gradient table, parse options,
// set_response(), init(), nSH(), and the same public members.  init() builds the FP64 problem
// matrices once on the host; the per-voxel fit (CSD::set / CSD::iterate, csd.h:236-269) runs on
// the GPU through the C ABI in include/mrcsd.h — see csd_engine.h.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "dense.h"
#include "gradient.h"

namespace MR {
namespace DWI {
namespace SDeconv {

// csd.h:29-35
constexpr double NORM_LAMBDA_MULTIPLIER = 0.0002;
constexpr uint32_t DEFAULT_CSD_LMAX = 8;
constexpr double DEFAULT_CSD_NEG_LAMBDA = 1.0;
constexpr double DEFAULT_CSD_NORM_LAMBDA = 1.0;
constexpr double DEFAULT_CSD_THRESHOLD = 0.0;
constexpr size_t DEFAULT_CSD_NITER = 50;

// The 300 built-in constraint directions [az el] (src/dwi/directions/predefined.cpp:92-393).
Matrix electrostatic_repulsion_300();

// Options of the CSD group (csd.cpp:28-56 plus -lmax / -directions of dwi2fod.cpp:38-56), as
// the command line would deliver them.  Unset fields keep the reference defaults.
struct CSDOptions {
  std::vector<uint32_t> lmax;        // -lmax (exactly one value for CSD)
  std::vector<double> filter;        // -filter
  Matrix directions;                 // -directions ([az el] rows)
  bool has_neg_lambda = false, has_norm_lambda = false, has_threshold = false, has_niter = false;
  double neg_lambda = 0, norm_lambda = 0, threshold = 0;
  long niter = 0;
  std::vector<double> shells;        // -shells
};

class CSD {
 public:
  class Shared {
   public:
    // csd.h:52-68: gradient table (already through get_DW_scheme), largest non-zero shell,
    // its direction matrix and lmax_data.
    explicit Shared(const Matrix& grad, const std::vector<double>& requested_shells = {});

    void parse_cmdline_options(const CSDOptions& opt);             // csd.h:71-99
    void set_response(const std::string& path);                    // csd.h:102-106
    void set_response(const std::vector<double>& coefs);           // csd.h:108-114
    void init();                                                   // csd.h:117-198
    size_t nSH() const { return size_t(HR_trans.cols); }           // csd.h:200-202

    Matrix grad;
    std::vector<double> response, init_filter;
    Matrix DW_dirs, HR_dirs;
    Matrix rconv, HR_trans, M, Mt_M;
    double neg_lambda, norm_lambda, threshold;
    std::vector<size_t> dwis;
    uint32_t lmax_response, lmax_data, lmax_cmdline, lmax;
    size_t niter;
  };

  // The per-voxel solver (csd.h:214-271) with the reference's call surface: set(), iterate()
  // until it returns true (at most shared.niter calls), FOD().  The fit itself runs on the GPU:
  // the first iterate() after set() solves the voxel there (one voxel per launch, up to niter
  // iterations) and the following calls replay its outcome, so that iterate() call c returns
  // true exactly when the reference's would (its negative set unchanged at call c) and FOD()
  // after call c is the reference's F after min(c, n) solves (recomputed on the GPU if asked for
  // mid-way).  Signals are taken as dwi2fod's load_data hands them over (dwi2fod.cpp:156-173):
  // single precision, negatives as 0, a non-finite value giving a zero FOD.  Copies (one per
  // thread, as ThreadedLoop makes them) share the device copy of the problem and own their device
  // buffers and HIP stream.  Batches of voxels belong to mrcsd_run (include/mrcsd.h), which has
  // the throughput; this class is the drop-in for code written against CSD.
  explicit CSD(const Shared& shared_data);
  CSD(const CSD& other);
  CSD& operator=(const CSD&) = delete;
  ~CSD();

  void set(const std::vector<double>& DW_signals);                // csd.h:236-243
  template <class V>
  void set(const V& DW_signals) {
    std::vector<double> b(size_t(DW_signals.size()));
    for (size_t i = 0; i < b.size(); ++i) b[i] = double(DW_signals[i]);
    set(b);
  }
  bool iterate();                                                 // csd.h:245-269
  const std::vector<double>& FOD() const;                         // csd.h:271

  const Shared& shared;

  struct Engine;     // device copy of the problem (shared by copies)
  struct Voxel;      // device buffers and stream of one object

 private:
  size_t solve(size_t budget) const;

  std::shared_ptr<Engine> engine_;
  mutable std::unique_ptr<Voxel> dev_;
  std::vector<float> b_;                // signals of the current voxel
  size_t calls_ = 0;                    // iterate() calls since set()
  mutable std::vector<double> F_;
  mutable size_t solved_budget_ = 0;    // niter of the cached GPU result (0: none)
  mutable size_t solved_n_ = 0;         // its loop counter (< solved_budget_: converged there)
  mutable size_t F_state_ = 0;          // solves that F_ reflects
};

// Plain-text numeric matrix reader with the rules of core/math/math.h:146-336: '#' starts a
// comment, values separated by whitespace, commas or semicolons, one matrix row per line.
Matrix load_matrix(const std::string& path);
std::vector<double> load_vector(const std::string& path);
Matrix parse_matrix_text(const std::string& text);

}  // namespace SDeconv
}  // namespace DWI
}  // namespace MR
