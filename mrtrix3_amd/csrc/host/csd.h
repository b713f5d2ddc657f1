// Host side of the single-tissue constrained spherical deconvolution path.
//
// MR::DWI::SDeconv::CSD::Shared keeps the call surface of the reference class
// (src/dwi/sdeconv/csd.h:49-212): construct from a gradient table, parse options,
// set_response(), init(), nSH(), and the same public members.  init() builds the FP64 problem
// matrices once on the host; the per-voxel fit (CSD::set / CSD::iterate, csd.h:236-269) runs on
// the GPU through the C ABI in include/mrcsd.h — see csd_engine.h.
#pragma once

#include <cstdint>
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
};

// Plain-text numeric matrix reader with the rules of core/math/math.h:146-336: '#' starts a
// comment, values separated by whitespace, commas or semicolons, one matrix row per line.
Matrix load_matrix(const std::string& path);
std::vector<double> load_vector(const std::string& path);
Matrix parse_matrix_text(const std::string& text);

}  // namespace SDeconv
}  // namespace DWI
}  // namespace MR
