"""Torch-tensor front end of the C ABI.

Names and argument order follow the reference's C++ entry points
(src/lu_nopiv_ll.hxx:588-793): ``lu_nopiv_ll_f32(handle, m, n, nb, d_a, ldda,
d_info)`` etc., where ``d_a`` is a column-major device tensor and ``d_info``
a one-element int32 device tensor. Errors raise RemifaError (a RuntimeError;
negative statuses name the invalid argument), as the reference raises
std::invalid_argument / std::runtime_error.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import (R_32F, R_16F, R_64F, R_16BF, R_8F_E4M3, FP32, FP16, TC32, TC16, BF16_TC32,
                   FP8_TC32, LL, RL, LuConfig, check, lib)

_DT = {torch.float32: R_32F, torch.float16: R_16F, torch.float64: R_64F,
       torch.bfloat16: R_16BF, torch.float8_e4m3fn: R_8F_E4M3}


def dtype_code(t) -> int:
    dt = t.dtype if isinstance(t, torch.Tensor) else t
    if dt not in _DT:
        raise TypeError(f"unsupported dtype {dt}")
    return _DT[dt]


def colmajor_empty(m: int, n: int, dtype=torch.float32, device="cuda", ld: int | None = None):
    """m x n column-major tensor view (stride (1, ld))."""
    ld = max(m, 1) if ld is None else ld
    return torch.empty((n, ld), dtype=dtype, device=device).t()[:m]


def to_colmajor(a: torch.Tensor, dtype=None, device=None):
    dtype = a.dtype if dtype is None else dtype
    device = a.device if device is None else device
    out = colmajor_empty(a.shape[0], a.shape[1], dtype=dtype, device=device)
    out.copy_(a)
    return out


def _ld(a: torch.Tensor) -> int:
    if a.dim() == 1:
        return max(a.shape[0], 1)
    if a.stride(0) != 1 and a.shape[0] > 1:
        raise ValueError("matrix must be column-major (stride(0) == 1)")
    if a.shape[1] <= 1:
        return max(a.stride(1), a.shape[0], 1)
    return a.stride(1)


def _p(t) -> C.c_void_p:
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


class Handle:
    """remifa_handle_t bound to a torch CUDA (HIP) stream."""

    def __init__(self, stream: torch.cuda.Stream | None = None, load_tuning: bool = True):
        self._h = C.c_void_p()
        s = torch.cuda.current_stream() if stream is None else stream
        self.stream = s
        check(lib().remifa_create(C.byref(self._h), C.c_void_p(s.cuda_stream)), "remifa_create")
        if load_tuning:
            from .autotune import load_batched_rules, load_rules
            rules = load_rules()
            if rules:
                self.set_gemm_rules(rules)
            brules = load_batched_rules()
            if brules:
                self.set_batched_tuning(brules)

    def set_batched_tuning(self, rules):
        """batched factorization shape per size bucket: list of [bucket,
        threads, workgroups per CU] (remifa_set_batched_tuning)"""
        flat = (C.c_int * (3 * len(rules)))(*[int(v) for r in rules for v in r])
        check(lib().remifa_set_batched_tuning(self._h, len(rules), C.cast(flat, C.c_void_p)),
              "remifa_set_batched_tuning")

    def set_gemm_rules(self, rules):
        """autotuned GEMM table: list of [kind, m_lo, m_hi, n_lo, n_hi, tile, split]"""
        flat = (C.c_int64 * (7 * len(rules)))(*[int(v) for r in rules for v in r])
        check(lib().remifa_set_gemm_rules(self._h, len(rules), C.cast(flat, C.c_void_p)),
              "remifa_set_gemm_rules")

    @property
    def value(self):
        return self._h

    def close(self):
        if self._h:
            lib().remifa_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream: torch.cuda.Stream):
        self.stream = stream
        check(lib().remifa_set_stream(self._h, C.c_void_p(stream.cuda_stream)), "remifa_set_stream")

    def set_workspace(self, ws: torch.Tensor | None):
        nbytes = 0 if ws is None else ws.numel() * ws.element_size()
        check(lib().remifa_set_workspace(self._h, _p(ws), nbytes), "remifa_set_workspace")
        self._ws = ws

    def set_tuning(self, panel_width: int = 0, split_k: int = 0):
        check(lib().remifa_set_tuning(self._h, panel_width, split_k), "remifa_set_tuning")

    def set_block_kernels(self, one_launch: bool = True, units_per_wave: int = 0,
                          wait_polls: int = 0):
        """LL block kernels: one launch (default) or diagonal then far kernel;
        units per helper wavefront and polls per wait of the former (0 = auto)"""
        check(lib().remifa_set_block_kernels(self._h, int(one_launch), units_per_wave,
                                             wait_polls), "remifa_set_block_kernels")

    def set_gemm_tile(self, tile: int = -1):
        check(lib().remifa_set_gemm_tile(self._h, tile), "remifa_set_gemm_tile")

    def set_lookahead(self, on: bool = True, reserve_cus: int = 0):
        """LL lookahead: the next outer block's update on a second stream,
        masked off `reserve_cus` CUs (0 = no mask)"""
        check(lib().remifa_set_lookahead(self._h, int(bool(on)), reserve_cus),
              "remifa_set_lookahead")

    # per-kernel timing -------------------------------------------------
    def timing(self, enable: bool = True):
        check(lib().remifa_timing_enable(self._h, int(enable)), "remifa_timing_enable")

    def timing_reset(self):
        check(lib().remifa_timing_reset(self._h), "remifa_timing_reset")

    def timing_query(self, cls: int):
        ms, n, w = C.c_double(), C.c_int64(), C.c_double()
        check(lib().remifa_timing_query(self._h, cls, C.byref(ms), C.byref(n), C.byref(w)),
              "remifa_timing_query")
        return ms.value, n.value, w.value

    def timing_records(self):
        """(cls, variant, ms, work) numpy arrays, one entry per recorded launch."""
        import numpy as np
        cnt = C.c_int64()
        check(lib().remifa_timing_records(self._h, 0, None, None, None, None, C.byref(cnt)),
              "remifa_timing_records")
        k = cnt.value
        cls, var = np.zeros(k, np.int32), np.zeros(k, np.int32)
        ms, work = np.zeros(k, np.float64), np.zeros(k, np.float64)
        vp = lambda a: C.c_void_p(a.ctypes.data)
        check(lib().remifa_timing_records(self._h, k, vp(cls), vp(var), vp(ms), vp(work),
                                          C.byref(cnt)), "remifa_timing_records")
        return cls, var, ms, work


# GEMM variant codes of the timer: opk << 8 | fp16-output << 4 | tile
_GEMM_OPK = {0: "fp16", 1: "bf16", 2: "fp32", 4: "bf16-stored"}
_GEMM_TILE = {0: (256, 32, 4, 1), 1: (32, 256, 1, 4), 2: (256, 64, 4, 1), 3: (64, 256, 1, 4),
              4: (256, 256, 2, 4), 5: (128, 128, 2, 2), 6: (256, 128, 2, 2), 7: (128, 256, 2, 2)}


def gemm_variant_name(v: int) -> str:
    """Template instantiation of gemm_mfma_kernel for a GEMM variant code."""
    opk, out16, tile = v >> 8, (v >> 4) & 15, v & 15
    bm, bn, wm, wn = _GEMM_TILE.get(tile, (0, 0, 0, 0))
    if opk == 3:
        return (f"gemm_fp8_kernel<{'_Float16' if out16 else 'float'}> "
                f"(scaled e4m3 operands, 128x128 tile)")
    if tile == 15:
        return f"splitk_reduce_kernel<{'_Float16' if out16 else 'float'}> (split-K partial sums)"
    if tile == 8:
        return (f"gemm_pair_kernel<{opk}, {'_Float16' if out16 else 'float'}> "
                f"({_GEMM_OPK.get(opk, '?')} operands, 256x32 + 32x256 tiles)")
    if tile == 4:
        return (f"gemm_big_kernel<{opk}, {'_Float16' if out16 else 'float'}, 32, 4, 8> "
                f"({_GEMM_OPK.get(opk, '?')} operands, 256x256 tile, LDS-DMA slabs)")
    return (f"gemm_mfma_kernel<{opk}, {'_Float16' if out16 else 'float'}, {bm}, {bn}, {wm}, {wn}>"
            f" ({_GEMM_OPK.get(opk, '?')} operands, {bm}x{bn} tile)")


def _h(handle) -> C.c_void_p:
    return handle.value if isinstance(handle, Handle) else handle


# ------------------------------------------------------------------ drivers
def lu_workspace_size(handle, algo: int, cfg: LuConfig, m: int, n: int, nb: int) -> int:
    out = C.c_size_t()
    check(lib().remifa_lu_workspace_size(_h(handle), algo, C.byref(cfg), m, n, nb, C.byref(out)),
          "remifa_lu_workspace_size")
    return out.value


def lu_nopiv(handle, a: torch.Tensor, nb: int, algo: int = LL, cfg: LuConfig | None = None,
             info: torch.Tensor | None = None, a_cpy: torch.Tensor | None = None, **kw) -> int:
    """In-place no-pivot LU of the square column-major tensor `a` (the
    generic lu_nopiv_ll / lu_nopiv_rl driver). Returns the info value (0, or
    the 1-based column of the first zero / non-finite pivot)."""
    if cfg is None:
        cfg = make_config(W=dtype_code(a), **kw)
    m, n = a.shape
    own_info = info is None
    if own_info:
        info = torch.zeros(1, dtype=torch.int32, device=a.device)
    ldc = _ld(a_cpy) if a_cpy is not None else 0
    check(lib().remifa_lu_nopiv(_h(handle), algo, C.byref(cfg), m, n, nb, _p(a), _ld(a), _p(info),
                                _p(a_cpy), ldc), "remifa_lu_nopiv")
    return int(info.item()) if own_info else 0


def make_config(W=R_32F, I=FP32, O=FP32, OP=None, IP=None, OF=None, IF=None, ib=8) -> LuConfig:
    OP = W if OP is None else OP
    IP = OP if IP is None else IP
    OF = OP if OF is None else OF
    IF = OF if IF is None else IF
    return LuConfig(W, I, O, OP, IP, OF, IF, ib)


def _helper(name, with_cpy):
    def f(handle, m, n, nb, d_a, ldda, d_info, d_a_f16=None, ld_a_f16=None):
        fn = getattr(lib(), "remifa_" + name)
        if with_cpy:
            ldc = ld_a_f16 if ld_a_f16 is not None else (_ld(d_a_f16) if d_a_f16 is not None else 0)
            st = fn(_h(handle), m, n, nb, _p(d_a), ldda, _p(d_info), _p(d_a_f16), ldc)
        else:
            st = fn(_h(handle), m, n, nb, _p(d_a), ldda, _p(d_info))
        check(st, name)
    f.__name__ = name
    f.__doc__ = f"remifa::{name} (src/lu_nopiv_ll.cxx) on the GPU."
    return f


# (name, takes the fp16 copy, W, I, O, OP, IP, OF, IF): src/lu_nopiv_ll.cxx
REFERENCE_HELPERS = [
    ("lu_nopiv_ll_f32", False, R_32F, FP32, FP32, R_32F, R_32F, R_32F, R_32F),
    ("lu_nopiv_ll_f32_fp32xtc32", True, R_32F, FP32, TC32, R_32F, R_32F, R_32F, R_32F),
    ("lu_nopiv_ll_f32_tc32xtc32", True, R_32F, TC32, TC32, R_32F, R_32F, R_32F, R_32F),
    ("lu_nopiv_ll_f16", False, R_16F, FP16, FP16, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_tc16xtc16", False, R_16F, TC16, TC16, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_tc32xtc32", False, R_16F, TC32, TC32, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_f16xtc32", False, R_16F, FP16, TC32, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_f16xtc32_fp32", False, R_16F, FP16, TC32, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_f32xt32", False, R_16F, FP32, TC32, R_32F, R_32F, R_32F, R_32F),
    ("lu_nopiv_ll_f16_f16xtc16", False, R_16F, FP16, TC16, R_16F, R_16F, R_16F, R_16F),
    ("lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32", False, R_16F, TC32, TC32, R_16F, R_32F,
     R_16F, R_32F),
    ("lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32", False, R_16F, TC32, TC32, R_32F, R_32F,
     R_32F, R_32F),
]

for _row in REFERENCE_HELPERS:
    globals()[_row[0]] = _helper(_row[0], _row[1])


# ------------------------------------------------------------------ kernels
def lu_nopiv_panel(handle, ib: int, m: int, n: int, d: torch.Tensor, l: torch.Tensor,
                   u: torch.Tensor, stat: torch.Tensor | None = None):
    """lu_nopiv_panel<T, ib> (src/lu_nopiv.cu:364-400) on column-major views."""
    check(lib().remifa_lu_nopiv_panel(_h(handle), dtype_code(d), ib, m, n, _p(d), _ld(d), _p(l),
                                      _ld(l), _p(u), _ld(u), _p(stat)), "remifa_lu_nopiv_panel")


def convert(handle, a: torch.Tensor, out: torch.Tensor):
    """convert<TA,TAO> (src/convert.cu:324-360): out <- a, rounded."""
    m, n = a.shape
    check(lib().remifa_convert(_h(handle), dtype_code(a), dtype_code(out), m, n, _p(a), _ld(a),
                               _p(out), _ld(out)), "remifa_convert")


def gemmex(handle, ctype: int, a: torch.Tensor, b: torch.Tensor, c: torch.Tensor,
           alpha: float = -1.0, beta: float = 1.0):
    """gemmex (src/gemm.cuh:24-40): c <- alpha a b + beta c."""
    m, k = a.shape
    n = b.shape[1]
    check(lib().remifa_gemmex(_h(handle), ctype, dtype_code(a), dtype_code(c), m, n, k, alpha,
                              _p(a), _ld(a), _p(b), _ld(b), beta, _p(c), _ld(c)), "remifa_gemmex")


def quantize_fp8(handle, a: torch.Tensor, rows: bool = True):
    """FP8 operand of a column-major f32/f16 matrix (remifa_quantize_fp8):
    rows=True quantizes the m x k matrix a by rows (A operand), rows=False the
    k x n matrix a by columns (B operand). Returns (codes, scales): uint8
    tensors [R, ldq] and [R, ldq // 32], row r holding the K-contiguous codes
    and the E8M0 scale bytes of row / column r (ldq = k rounded up to 128)."""
    r, k = (a.shape[0], a.shape[1]) if rows else (a.shape[1], a.shape[0])
    ldq = max(-(-k // 128) * 128, 128)
    q = torch.zeros((r, ldq), dtype=torch.uint8, device=a.device)
    sc = torch.full((r, ldq // 32), 127, dtype=torch.uint8, device=a.device)
    check(lib().remifa_quantize_fp8(_h(handle), dtype_code(a), int(rows), r, k, _p(a), _ld(a),
                                    _p(q), ldq, _p(sc), ldq // 32), "remifa_quantize_fp8")
    return q, sc


def gemm_fp8(handle, qa: torch.Tensor, sa: torch.Tensor, qb: torch.Tensor, sb: torch.Tensor,
             k: int, c: torch.Tensor, alpha: float = -1.0, beta: float = 1.0):
    """c <- alpha A B + beta c from quantize_fp8 operands (remifa_gemm_fp8)."""
    m, n = c.shape
    check(lib().remifa_gemm_fp8(_h(handle), dtype_code(c), m, n, k, alpha, _p(qa), qa.stride(0),
                                _p(sa), sa.stride(0), _p(qb), qb.stride(0), _p(sb), sb.stride(0),
                                beta, _p(c), _ld(c)), "remifa_gemm_fp8")


# ------------------------------------------------------------------ solves
def getrs_nopiv_noalloc(handle, lu: torch.Tensor, b: torch.Tensor, trans: bool = False):
    """getrs_nopiv with its scratch in a static device array (no allocation;
    n <= 32768; calls that overlap on one GPU race)."""
    n = lu.shape[0]
    nrhs = b.shape[1] if b.dim() == 2 else 1
    check(lib().remifa_getrs_nopiv_noalloc(_h(handle), dtype_code(lu), int(trans), n, nrhs,
                                           _p(lu), _ld(lu), _p(b), _ld(b)),
          "remifa_getrs_nopiv_noalloc")


def getrs_nopiv(handle, lu: torch.Tensor, b: torch.Tensor, trans: bool = False):
    """b (fp32, n x nrhs column-major) <- (LU)^-1 b or (LU)^-T b."""
    n = lu.shape[0]
    nrhs = b.shape[1] if b.dim() == 2 else 1
    check(lib().remifa_getrs_nopiv(_h(handle), dtype_code(lu), int(trans), n, nrhs, _p(lu),
                                   _ld(lu), _p(b), _ld(b)), "remifa_getrs_nopiv")


def residual_f64(handle, a: torch.Tensor, x: torch.Tensor, b: torch.Tensor, r: torch.Tensor):
    n = a.shape[0]
    nrhs = b.shape[1] if b.dim() == 2 else 1
    check(lib().remifa_residual_f64(_h(handle), n, nrhs, _p(a), _ld(a), _p(x), _ld(x), _p(b),
                                    _ld(b), _p(r), _ld(r)), "remifa_residual_f64")


def ir_solve(handle, a: torch.Tensor, lu: torch.Tensor, b: torch.Tensor, x: torch.Tensor,
             tol: float = 1e-14, maxit: int = 30, trans: bool = False):
    """Iterative refinement (FP64 residual) of A X = B, or A^T X = B when
    trans. Returns (iters, backward_error, history); iters is -1 if tol was
    not reached."""
    n = a.shape[0]
    nrhs = b.shape[1] if b.dim() == 2 else 1
    it, be = C.c_int(), C.c_double()
    hist = (C.c_double * (maxit + 1))()
    fn = "remifa_ir_solve_trans" if trans else "remifa_ir_solve"
    check(getattr(lib(), fn)(_h(handle), dtype_code(lu), n, nrhs, _p(a), _ld(a), _p(lu),
                             _ld(lu), _p(b), _ld(b), _p(x), _ld(x), tol, maxit, C.byref(it),
                             C.byref(be), C.cast(hist, C.c_void_p)), fn)
    k = it.value if it.value >= 0 else maxit
    return it.value, be.value, list(hist)[: k + 1]


# ------------------------------------------------------------------ error measures
def error_measures(handle, a: torch.Tensor, x: torch.Tensor, b: torch.Tensor,
                   lu: torch.Tensor | None = None) -> dict:
    """tests/errors.hxx measures of the solutions x (FP64) of A x = b on the
    GPU: normwise backward error, componentwise error with |L||U| (needs the
    factors lu), Oettli-Prager form."""
    n = a.shape[0]
    nrhs = b.shape[1] if b.dim() == 2 else 1
    out = (C.c_double * 3)()
    lu_t = dtype_code(lu) if lu is not None else R_32F
    check(lib().remifa_error_measures(_h(handle), lu_t, n, nrhs, _p(a), _ld(a), _p(lu),
                                      _ld(lu) if lu is not None else n, _p(b), _ld(b), _p(x),
                                      _ld(x), out), "remifa_error_measures")
    return {"nwerr": out[0], "cwerr": out[1], "cwerr_ax": out[2]}


def factor_error(handle, a: torch.Tensor, lu: torch.Tensor) -> dict:
    """tests/errors.hxx:251-346 factor error of the in-place factors lu of
    a, on the GPU in FP64."""
    n = a.shape[0]
    out = (C.c_double * 4)()
    check(lib().remifa_factor_error(_h(handle), dtype_code(lu), n, _p(a), _ld(a), _p(lu),
                                    _ld(lu), out), "remifa_factor_error")
    return {"factor_cwerr": out[0], "factor_abs": out[1], "factor_rel_fro": out[2],
            "factor_fro": out[3]}


# ------------------------------------------------------------------ generators
def gen_diagdom(handle, a: torch.Tensor, seed: int = 1, alpha: float = 0.0):
    """Fill the n x n fp32 column-major `a`: uniform(-1,1) off-diagonal from the
    seeded minstd_rand0 stream (row-major draws), a_ii = alpha * sum_j!=i |a_ij|
    (alpha <= 0: gen_diagdom of tests/genmat.hxx, a_ii = n, default seed)."""
    n = a.shape[0]
    check(lib().remifa_gen_diagdom(_h(handle), n, seed, alpha, _p(a), _ld(a)), "remifa_gen_diagdom")


def gen_randmat(handle, a: torch.Tensor, seed: int = 1):
    """gen_randmat (tests/common.hxx:298-319): uniform(-1,1), column-major
    draws, into the fp32 column-major `a`."""
    m, n = a.shape
    check(lib().remifa_gen_randmat(_h(handle), m, n, seed, _p(a), _ld(a)), "remifa_gen_randmat")


def gen_arrowhead(handle, a: torch.Tensor, lo: float = 0.0, hi: float = 1.0):
    """gen_arrowhead (tests/common.hxx:391-455, k = 256) into the square
    fp32 column-major `a`."""
    n = a.shape[0]
    check(lib().remifa_gen_arrowhead(_h(handle), n, lo, hi, _p(a), _ld(a)),
          "remifa_gen_arrowhead")


def geequb(handle, a: torch.Tensor, r: torch.Tensor, c: torch.Tensor) -> dict:
    """LAPACK sgeequb on the GPU: power-of-two scalings into r (m) and c (n);
    returns rowcnd, colcnd, amax, info."""
    m, n = a.shape
    rc, cc, am, info = C.c_float(), C.c_float(), C.c_float(), C.c_int()
    check(lib().remifa_geequb(_h(handle), m, n, _p(a), _ld(a), _p(r), _p(c), C.byref(rc),
                              C.byref(cc), C.byref(am), C.byref(info)), "remifa_geequb")
    return {"rowcnd": rc.value, "colcnd": cc.value, "amax": am.value, "info": info.value}


def laqge(handle, a: torch.Tensor, r: torch.Tensor, c: torch.Tensor, rowcnd: float,
          colcnd: float, amax: float) -> str:
    """LAPACK slaqge on the GPU: scales `a` in place; returns EQUED."""
    m, n = a.shape
    eq = C.create_string_buffer(1)
    check(lib().remifa_laqge(_h(handle), m, n, _p(a), _ld(a), _p(r), _p(c), rowcnd, colcnd,
                             amax, eq), "remifa_laqge")
    return eq.raw.decode()


def conv_error(handle, a: torch.Tensor, conv: torch.Tensor) -> dict:
    """--check-conv (tests/common.hxx:107-174): relative and absolute
    conversion error of the copy `conv` (fp16 / bf16 / fp32) of `a`."""
    m, n = a.shape
    out = (C.c_double * 2)()
    check(lib().remifa_conv_error(_h(handle), dtype_code(conv), m, n, _p(a), _ld(a), _p(conv),
                                  _ld(conv), out), "remifa_conv_error")
    return {"rel": out[0], "abs": out[1]}


def gen_uniform(handle, b: torch.Tensor, seed: int):
    m, n = b.shape
    check(lib().remifa_gen_uniform(_h(handle), m, n, seed, _p(b), _ld(b)), "remifa_gen_uniform")


# ------------------------------------------------------------------ batched
def lu_nopiv_batched(handle, ctype: int, n: torch.Tensor, ld: torch.Tensor, ptrs: torch.Tensor,
                     info: torch.Tensor, max_n: int):
    check(lib().remifa_lu_nopiv_batched(_h(handle), ctype, n.numel(), _p(n), _p(ld), _p(ptrs),
                                        _p(info), max_n), "remifa_lu_nopiv_batched")


def getrs_nopiv_batched(handle, n: torch.Tensor, ld: torch.Tensor, ptrs: torch.Tensor,
                        bptrs: torch.Tensor, max_n: int):
    check(lib().remifa_getrs_nopiv_batched(_h(handle), n.numel(), _p(n), _p(ld), _p(ptrs),
                                           _p(bptrs), max_n), "remifa_getrs_nopiv_batched")
