"""MI355X-native spectral-quadrature (SQ) engine for SQMiniApp's per-node Lanczos hot path.

The compute path is libsqamd.so (HIP kernels for gfx950 + C ABI, include/sq_amd.h);
this package binds it and provides deck generators for the benchmark supercells.
"""
from ._lib import (SQ_FP32, SQ_FP64, SQ_NL_KB, SQ_NL_REF_COLLAPSED, Context, Mlp, SQError, lib, tune_select)  # noqa: F401

__all__ = ["Context", "SQError", "lib", "SQ_FP32", "SQ_FP64", "SQ_NL_KB", "SQ_NL_REF_COLLAPSED"]
