// Host side of multi-shell multi-tissue CSD: MR::DWI::SDeconv::MSMT_CSD::Shared keeps the call
// surface of src/dwi/sdeconv/msmt_csd.h:47-257 and builds the ICLS problem once; the per-voxel
// solve runs on the GPU (kernels/icls_solve.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dense.h"
#include "gradient.h"
#include "icls.h"

namespace MR {
namespace DWI {
namespace SDeconv {

constexpr uint32_t DEFAULT_MSMTCSD_LMAX = 8;           // msmt_csd.h:31
constexpr double DEFAULT_MSMTCSD_NORM_LAMBDA = 1.0e-10;  // msmt_csd.h:32
constexpr double DEFAULT_MSMTCSD_NEG_LAMBDA = 1.0e-10;   // msmt_csd.h:33

class MSMT_CSD {
 public:
  class Shared {
   public:
    // msmt_csd.h:49-54; requested_shells: the b-values of a -shells option (empty: all shells)
    explicit Shared(const Matrix& grad, const std::vector<double>& requested_shells = {});
    void set_responses(const std::vector<Matrix>& matrices);      // msmt_csd.h:91-95
    void set_responses(const std::vector<std::string>& files);    // msmt_csd.h:76-89
    void init();                                                  // msmt_csd.h:100-218
    size_t num_shells() const { return shells.count(); }
    size_t num_tissues() const { return responses.size(); }

    Matrix grad;
    Shells shells;
    Matrix HR_dirs;
    std::vector<uint32_t> lmax, lmax_response;
    std::vector<Matrix> responses;
    std::vector<std::string> response_files;
    Math::ICLS::Problem problem;
    Matrix C, A;   // problem and constraint matrices handed to the ICLS problem
    double solution_min_norm_regularisation, constraint_min_norm_regularisation;

   private:
    void prepare_responses();                                     // msmt_csd.h:237-255
  };

  // msmt_csd.h:262-275: the per-voxel solve over all volumes of a voxel (ICLS::Solver on the GPU).
  explicit MSMT_CSD(const Shared& shared_data) : niter(0), shared(shared_data), solver(shared_data.problem) {}
  void operator()(const std::vector<double>& data, std::vector<double>& output) { niter = solver(output, data); }

  size_t niter;
  const Shared& shared;

 private:
  Math::ICLS::Solver solver;
};

}  // namespace SDeconv
}  // namespace DWI
}  // namespace MR
