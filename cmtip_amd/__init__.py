"""cmtip_amd: the cartesian-MTIP hot path (apeck12/cmtip) as HIP kernels for MI355X.

Subpackages mirror the reference's modules: ``nufft`` (type-2 forward / type-1
adjoint), ``alignment`` (slicing, matching, errors) and ``autocorrelation`` (the W x
= d solve).  The kernels live in ``libcmtip_amd.so`` (C ABI in include/cmtip_amd.h),
built by ``python -m cmtip_amd.build``.
"""
__version__ = "0.1.0"
