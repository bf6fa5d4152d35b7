"""rgf_amd: the recursive Green's function S-matrix solve of chiaweihsu/RGF on MI355X.

The numerical work lives in ``librgf.so`` (hand-written gfx950 HIP kernels
behind the C ABI of ``include/rgf.h``); this package mirrors the reference's
MATLAB functions on top of it.
"""
from ._lib import RGFError, load_library, trim  # noqa: F401
from .api import (  # noqa: F401
    build_laplacian_1d,
    cal_smatrix_RGF,
    cal_smatrix_RGF_3D,
    cal_smatrix_RGF_batched,
    dense_gemm,
    dense_inverse,
    setup_channels,
    setup_channels_3D,
    transmission_gradient,
)
from .sweep import load_scan, save_scan, scan  # noqa: F401

__all__ = [
    "cal_smatrix_RGF", "cal_smatrix_RGF_3D", "setup_channels", "setup_channels_3D",
    "build_laplacian_1d", "load_library", "RGFError", "dense_inverse", "dense_gemm",
    "transmission_gradient", "cal_smatrix_RGF_batched", "scan", "save_scan", "load_scan", "trim",
]
