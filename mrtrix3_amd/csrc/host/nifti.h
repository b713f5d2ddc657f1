// NIfTI-1 image input for the commands (core/formats/nifti1.cpp, nifti1_gz.cpp and
// core/file/nifti_utils.cpp:83-360 in the reference): .nii and .nii.gz, read into the same
// in-memory image as a .mif.
#pragma once

#include <string>

#include "mif.h"

namespace mrcsd {

// Header fields as the reference interprets them: dimensions, voxel sizes, the data type in either
// byte order with scl_slope / scl_inter, and the transform from the s-form when its code is set,
// else the q-form (quaternion, qfac), each axis normalised to unit length (nifti_utils.cpp:240-330).
MifImage read_nifti(const std::string& path);

// read_nifti for .nii / .nii.gz, read_mif otherwise.
MifImage read_image(const std::string& path);

}  // namespace mrcsd
