"""MI355X-native constrained spherical deconvolution (the dwi2fod csd / msmt_csd hot path).

Native library: mrtrix3_amd/lib/libmrcsd.so (HIP kernels for gfx950 + C ABI, include/mrcsd.h),
built in-tree by mrtrix3_amd/build.py.  Python modules:
  csd      CSD::Shared / CSD mirror over the C ABI
  phantom  seeded synthetic workloads
"""
__version__ = "0.1"
