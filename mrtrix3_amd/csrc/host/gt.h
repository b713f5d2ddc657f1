// Host set-up of global tractography's external energy (MR::DWI::Tractography::GT::
// ExternalEnergyComputer, src/dwi/tractography/GT/externalenergy.cpp:30-112): the matrices that
// map a voxel's track orientation distribution (TOD, SH coefficients) and isotropic fractions to its
// DW signal.  The per-voxel energies (calcEnergy, :245-252, an NNLS fit per voxel) run on the GPU.
#pragma once

#include <vector>

#include "dense.h"

namespace MR {
namespace DWI {
namespace Tractography {
namespace GT {

using mrcsd::Matrix;

// externalenergy.cpp:55-101: K [nvol x NforL(lmax)] (row r: the SH delta of volume r's direction
// convolved with the white-matter response of its shell, times `weight`) and Ak [nvol x (nf + 1)]
// (column 0: the white-matter response's l = 0 term / sqrt(4 pi); column j + 1: isotropic response j
// on the volume's shell).  grad: the scheme as get_DW_scheme returns it; resp_wm [nshell x ncoef]
// zonal coefficients per shell; resp_iso[j] one value per shell.
void external_energy_kernels(const Matrix& grad, const Matrix& resp_wm, const std::vector<std::vector<double>>& resp_iso,
                             int lmax, double weight, Matrix& K, Matrix& Ak);

// Math::SH::delta (core/math/SH.h:279-303): the SH basis at a unit direction, up to lmax.
std::vector<double> sh_delta(const double* unit_dir, int lmax);

}  // namespace GT
}  // namespace Tractography
}  // namespace DWI
}  // namespace MR
