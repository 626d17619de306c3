"""remifa_amd: MI355X (gfx950) no-pivot LU factorization.

Python view of ``libremifa_hip.so`` (C ABI declared in ``include/remifa.h``).
Device memory and streams come from PyTorch (ROCm); every computation runs in
the library's HIP kernels. There is no CPU fallback: importing this package
without the built library raises ``ImportError``.

Matrices are column-major, as in the reference (flipflapflop/remifa). A
column-major m x n matrix with leading dimension ld is a torch tensor view
with ``stride() == (1, ld)``, e.g. ``torch.empty(n, ld).t()[:m]``; helpers
``colmajor_empty`` / ``to_colmajor`` build such views.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _lib
from ._lib import (  # noqa: F401  (re-exported enums)
    R_32F, R_16F, R_64F, R_16BF, R_8F_E4M3,
    FP32, FP16, TC32, TC16, BF16_TC32, FP8_TC32,
    LL, RL, RemifaError, LuConfig, lib, library_path,
)
from .api import (  # noqa: F401
    Handle, colmajor_empty, to_colmajor, dtype_code,
    lu_nopiv, lu_nopiv_panel, convert, gemmex, quantize_fp8, gemm_fp8, getrs_nopiv, residual_f64,
    ir_solve, error_measures, factor_error, gen_randmat, gen_arrowhead, geequb, laqge,
    conv_error, getrs_nopiv_noalloc,
    gen_diagdom, gen_uniform, lu_workspace_size, lu_nopiv_batched, getrs_nopiv_batched,
    REFERENCE_HELPERS,
)
from .api import (  # noqa: F401  (the reference's entry-point names)
    lu_nopiv_ll_f32, lu_nopiv_ll_f32_fp32xtc32, lu_nopiv_ll_f32_tc32xtc32, lu_nopiv_ll_f16,
    lu_nopiv_ll_f16_tc16xtc16, lu_nopiv_ll_f16_tc32xtc32, lu_nopiv_ll_f16_f16xtc32,
    lu_nopiv_ll_f16_f16xtc32_fp32, lu_nopiv_ll_f16_f32xt32, lu_nopiv_ll_f16_f16xtc16,
    lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32,
    lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32,
)

from . import autograd  # noqa: F401,E402  (torch.autograd.Function X = A^-1 B)

__all__ = [n for n in dir() if not n.startswith("_")]
