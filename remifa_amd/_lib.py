"""ctypes declarations of libremifa_hip.so (include/remifa.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libremifa_hip.so")

R_32F, R_16F, R_64F, R_16BF, R_8F_E4M3 = 0, 1, 2, 3, 4
FP32, FP16, TC32, TC16, BF16_TC32, FP8_TC32 = 0, 1, 2, 3, 4, 5
LL, RL = 0, 1

ERR_HIP, ERR_ALLOC, ERR_UNSUPPORTED, ERR_WORKSPACE = 1000, 1001, 1002, 1003


class RemifaError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str = ""):
        self.status = status
        if status < 0:
            text = f"{func}: argument {-status} is invalid"
        else:
            text = f"{func}: error {status}" + (f" ({msg})" if msg else "")
        super().__init__(text)


class LuConfig(C.Structure):
    """remifa_lu_config_t: the precision tuple of lu_nopiv_ll<W,I,O,ib,OP,IP,OF,IF>."""
    _fields_ = [("W", C.c_int), ("I", C.c_int), ("O", C.c_int), ("OP", C.c_int),
                ("IP", C.c_int), ("OF", C.c_int), ("IF", C.c_int), ("ib", C.c_int64)]

    def __repr__(self):
        return (f"LuConfig(W={self.W}, I={self.I}, O={self.O}, OP={self.OP}, IP={self.IP}, "
                f"OF={self.OF}, IF={self.IF}, ib={self.ib})")


_lib = None

_i, _i64, _u64, _vp, _d, _f, _sz = (C.c_int, C.c_int64, C.c_uint64, C.c_void_p, C.c_double,
                                    C.c_float, C.c_size_t)
_pi, _pd, _pi64 = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)

SIGNATURES = {
    "remifa_create": [C.POINTER(_vp), _vp],
    "remifa_destroy": [_vp],
    "remifa_set_stream": [_vp, _vp],
    "remifa_get_stream": [_vp, C.POINTER(_vp)],
    "remifa_set_workspace": [_vp, _vp, _sz],
    "remifa_lu_workspace_size": [_vp, _i, C.POINTER(LuConfig), _i64, _i64, _i64, C.POINTER(_sz)],
    "remifa_set_tuning": [_vp, _i, _i],
    "remifa_set_block_kernels": [_vp, _i, _i, _i64],
    "remifa_set_gemm_tile": [_vp, _i],
    "remifa_set_lookahead": [_vp, _i, _i],
    "remifa_get_lookahead": [_vp, _pi, _pi],
    "remifa_last_error": [],
    "remifa_timing_enable": [_vp, _i],
    "remifa_timing_query": [_vp, _i, _pd, _pi64, _pd],
    "remifa_timing_reset": [_vp],
    "remifa_timing_records": [_vp, _i64, _vp, _vp, _vp, _vp, _pi64],
    "remifa_lu_nopiv": [_vp, _i, C.POINTER(LuConfig), _i64, _i64, _i64, _vp, _i64, _vp, _vp, _i64],
    "remifa_lu_nopiv_ll_f32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f32_fp32xtc32": [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _i],
    "remifa_lu_nopiv_ll_f32_tc32xtc32": [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _i],
    "remifa_lu_nopiv_ll_f16": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_tc16xtc16": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_tc32xtc32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_f16xtc32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_f16xtc32_fp32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_f32xt32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_f16xtc16": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32": [_vp, _i, _i, _i, _vp, _i, _vp],
    "remifa_lu_nopiv_panel": [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp],
    "remifa_convert": [_vp, _i, _i, _i64, _i64, _vp, _i64, _vp, _i64],
    "remifa_gemmex": [_vp, _i, _i, _i, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i],
    "remifa_quantize_fp8": [_vp, _i, _i, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64],
    "remifa_gemm_fp8": [_vp, _i, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                        _f, _vp, _i64],
    "remifa_getrs_nopiv": [_vp, _i, _i, _i64, _i64, _vp, _i64, _vp, _i64],
    "remifa_getrs_nopiv_noalloc": [_vp, _i, _i, _i64, _i64, _vp, _i64, _vp, _i64],
    "remifa_residual_f64": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64],
    "remifa_ir_solve": [_vp, _i, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _d, _i,
                        _pi, _pd, _vp],
    "remifa_ir_solve_trans": [_vp, _i, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _d,
                              _i, _pi, _pd, _vp],
    "remifa_set_gemm_rules": [_vp, _i, _vp],
    "remifa_error_measures": [_vp, _i, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                              _pd],
    "remifa_factor_error": [_vp, _i, _i64, _vp, _i64, _vp, _i64, _pd],
    "remifa_gen_diagdom": [_vp, _i64, _u64, _d, _vp, _i64],
    "remifa_gen_uniform": [_vp, _i64, _i64, _u64, _vp, _i64],
    "remifa_gen_randmat": [_vp, _i64, _i64, _u64, _vp, _i64],
    "remifa_gen_arrowhead": [_vp, _i64, _f, _f, _vp, _i64],
    "remifa_geequb": [_vp, _i64, _i64, _vp, _i64, _vp, _vp, C.POINTER(_f), C.POINTER(_f),
                      C.POINTER(_f), _pi],
    "remifa_laqge": [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _f, _f, _f, C.c_char_p],
    "remifa_conv_error": [_vp, _i, _i64, _i64, _vp, _i64, _vp, _i64, _pd],
    "remifa_lu_nopiv_batched": [_vp, _i, _i64, _vp, _vp, _vp, _vp, _i],
    "remifa_getrs_nopiv_batched": [_vp, _i64, _vp, _vp, _vp, _vp, _i],
    "remifa_set_batched_tuning": [_vp, _i, _vp],
}


def lib():
    """Load libremifa_hip.so; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(
                f"{library_path} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (or python remifa_amd/build.py). There is no CPU fallback.")
        L = C.CDLL(library_path)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "remifa_last_error" else C.c_int
        _lib = L
    return _lib


def check(status: int, func: str) -> None:
    if status != 0:
        msg = lib().remifa_last_error().decode(errors="replace") if status > 0 else ""
        raise RemifaError(status, func, msg)
