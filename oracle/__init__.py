"""CPU oracle for the no-pivot LU hot path of flipflapflop/remifa.

TEST INFRASTRUCTURE ONLY. This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product (``remifa_amd``) never imports it and never falls back
to it.

The numerics live in ``remifa_oracle.c`` (a restatement of the reference's
``src/lu_nopiv*.{cu,hxx}``, ``src/convert.cu``, ``src/gemm.cu`` and
``tests/{errors,genmat}.hxx``); this module wraps it with numpy arrays.
Matrices are column-major (Fortran order) like the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libremifa_oracle.so")

F32, F16, F64, BF16, E4M3 = 0, 1, 2, 3, 4
FP32, FP16, TC32, TC16, BF16TC32, FP8TC32 = 0, 1, 2, 3, 4, 5

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force and os.path.exists(_SO):
        os.remove(_SO)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = C.CDLL(_SO)
        _declare(_lib)
    return _lib


class LuCfg(C.Structure):
    _fields_ = [("W", C.c_int), ("I", C.c_int), ("O", C.c_int),
                ("OP", C.c_int), ("IP", C.c_int), ("OF", C.c_int),
                ("IF", C.c_int), ("ib", C.c_int64)]


def _declare(L):
    i64, u64, vp, dp = C.c_int64, C.c_uint64, C.c_void_p, C.c_double
    sig = {
        "orc_minstd_seed": (u64, [u64]),
        "orc_minstd_jump": (u64, [u64, u64]),
        "orc_minstd_uniform": (C.c_float, [u64]),
        "orc_minstd_stream": (None, [u64, i64, vp]),
        "orc_gen_diagdom": (None, [i64, i64, vp, i64]),
        "orc_gen_diagdom_alpha": (None, [i64, u64, dp, vp, i64]),
        "orc_gen_uniform": (None, [i64, i64, u64, vp, i64]),
        "orc_gen_randmat": (None, [i64, i64, u64, vp, i64]),
        "orc_gen_arrowhead": (None, [i64, i64, C.c_float, C.c_float, vp, i64]),
        "orc_geequb": (C.c_int, [i64, i64, vp, i64, vp, vp, vp, vp, vp]),
        "orc_laqge": (C.c_char, [i64, i64, vp, i64, vp, vp, C.c_float, C.c_float, C.c_float]),
        "orc_f32_to_f16": (C.c_uint16, [C.c_float]),
        "orc_f16_to_f32": (C.c_float, [C.c_uint16]),
        "orc_f64_to_f16": (C.c_uint16, [dp]),
        "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
        "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
        "orc_f32_to_e4m3": (C.c_uint8, [C.c_float]),
        "orc_e4m3_to_f32": (C.c_float, [C.c_uint8]),
        "orc_round": (C.c_float, [C.c_int, C.c_float]),
        "orc_convert": (C.c_int, [C.c_int, C.c_int, i64, i64, vp, i64, vp, i64]),
        "orc_lu_nopiv_panel": (C.c_int, [C.c_int, i64, i64, i64, vp, i64, vp, i64,
                                         vp, i64, vp]),
        "orc_quantise_e4m3_codes": (None, [i64, i64, vp, i64, vp, vp, C.c_int]),
        "orc_set_sum_order": (None, [C.c_int]),
        "orc_gemm_sub": (None, [C.c_int, i64, i64, i64, vp, i64, vp, i64, vp, i64,
                                C.c_int]),
        "orc_lu_nopiv_ll_typed": (C.c_int, [C.POINTER(LuCfg), i64, i64, i64, vp, i64]),
        "orc_lu_nopiv_rl_typed": (C.c_int, [C.POINTER(LuCfg), i64, i64, i64, vp, i64]),
        "orc_lu_f64": (C.c_int, [i64, vp, i64]),
        "orc_getrs_nopiv_f32": (None, [i64, i64, vp, i64, vp, i64]),
        "orc_getrs_nopiv_f64": (None, [i64, i64, vp, i64, vp, i64]),
        "orc_getrs_nopiv_trans_f64": (None, [i64, i64, vp, i64, vp, i64]),
        "orc_unsym_backward_error": (dp, [i64, i64, vp, i64, vp, i64, i64, vp, i64]),
        "orc_unsym_compwise_error": (dp, [i64, i64, vp, i64, vp, vp, vp, i64, vp]),
        "orc_factor_backward_error": (dp, [i64, vp, i64, vp, i64]),
        "orc_residual_f64": (None, [i64, i64, vp, i64, vp, i64, vp, i64, vp, i64]),
        "orc_ir_solve": (C.c_int, [i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, dp,
                                   C.c_int, vp]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def set_sum_order(order: int) -> None:
    """0: k ascending (the restatement); 1: k descending (fp32 sums only)"""
    lib().orc_set_sum_order(int(order))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ---------------------------------------------------------------- generators
def minstd_stream(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_minstd_stream(seed, count, _p(out))
    return out


def gen_diagdom(m: int, n: int | None = None) -> np.ndarray:
    """tests/genmat.hxx:68-102 (default-seeded engine, a_ii = m)."""
    n = m if n is None else n
    a = np.zeros((m, n), np.float32, order="F")
    lib().orc_gen_diagdom(m, n, _p(a), m)
    return a


def gen_diagdom_alpha(n: int, seed: int, alpha: float) -> np.ndarray:
    a = np.zeros((n, n), np.float32, order="F")
    lib().orc_gen_diagdom_alpha(n, seed, float(alpha), _p(a), n)
    return a


def gen_uniform(m: int, n: int, seed: int) -> np.ndarray:
    b = np.zeros((m, n), np.float32, order="F")
    lib().orc_gen_uniform(m, n, seed, _p(b), m)
    return b


# ---------------------------------------------------------------- conversions
_NP = {F32: np.float32, F64: np.float64, F16: np.uint16, BF16: np.uint16,
       E4M3: np.uint8}


def gen_randmat(m: int, n: int, seed: int = 1) -> np.ndarray:
    """tests/common.hxx:298-319 (column-major draws)"""
    a = np.zeros((m, n), np.float32, order="F")
    lib().orc_gen_randmat(m, n, seed, _p(a), m)
    return a


def gen_arrowhead(n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """tests/common.hxx:391-455 (k = 256)"""
    a = np.zeros((n, n), np.float32, order="F")
    lib().orc_gen_arrowhead(n, n, lo, hi, _p(a), n)
    return a


def geequb(a):
    """LAPACK sgeequb restated: (r, c, rowcnd, colcnd, amax, info)"""
    a = _fortran(a, np.float32)
    m, n = a.shape
    r, c = np.zeros(m, np.float32), np.zeros(n, np.float32)
    out = np.zeros(3, np.float32)
    info = lib().orc_geequb(m, n, _p(a), m, _p(r), _p(c), _p(out[0:]), _p(out[1:]),
                            _p(out[2:]))
    return r, c, float(out[0]), float(out[1]), float(out[2]), int(info)


def laqge(a, r, c, rowcnd, colcnd, amax):
    """LAPACK slaqge restated: (scaled copy of a, equed)"""
    a = np.array(_fortran(a, np.float32), order="F", copy=True)
    m, n = a.shape
    r = np.ascontiguousarray(r, np.float32)
    c = np.ascontiguousarray(c, np.float32)
    eq = lib().orc_laqge(m, n, _p(a), m, _p(r), _p(c), rowcnd, colcnd, amax)
    return a, eq.decode()


def convert(a: np.ndarray, src: int, dst: int) -> np.ndarray:
    a = _fortran(a, _NP[src])
    m, n = a.shape
    b = np.zeros((m, n), _NP[dst], order="F")
    st = lib().orc_convert(src, dst, m, n, _p(a), max(m, 1), _p(b), max(m, 1))
    if st:
        raise ValueError(f"orc_convert failed: {st}")
    return b


def f32_to_f16_bits(a) -> np.ndarray:
    return convert(np.atleast_2d(np.asarray(a, np.float32)), F32, F16)


def f16_bits_to_f32(h) -> np.ndarray:
    return convert(np.atleast_2d(np.asarray(h, np.uint16)), F16, F32)


def round_to(prec: int, a) -> np.ndarray:
    """Round float32 values to `prec` and return them as float32."""
    a = np.atleast_2d(np.asarray(a, np.float32))
    if prec == F32:
        return a.copy()
    return convert(convert(a, F32, prec), prec, F32)


# ---------------------------------------------------------------- panel/gemm
def lu_nopiv_panel(prec: int, ib: int, m: int, n: int, d, l, u):
    """src/lu_nopiv.cu:364-400. d: n x n snapshot, l: m x n (in/out),
    u: n x m (in/out; columns n..m-1 are solved). Arrays in storage dtype."""
    dt = _NP[prec]
    d = _fortran(d, dt)
    l = _fortran(l, dt).copy(order="F")
    u = _fortran(u, dt).copy(order="F")
    stat = np.zeros(1, np.int32)
    st = lib().orc_lu_nopiv_panel(prec, ib, m, n, _p(d), max(d.shape[0], 1), _p(l),
                                  max(l.shape[0], 1), _p(u), max(u.shape[0], 1),
                                  _p(stat))
    if st:
        raise ValueError(f"orc_lu_nopiv_panel: bad argument {-st}")
    return l, u, int(stat[0])


def quantise_e4m3(a, rows: bool = True):
    """Codes and scale exponents of the FP8TC32 operand format: a is m x k
    quantised by rows (rows=True) or k x m by columns (rows=False).
    Returns codes uint8 [r, k] and exponents int32 [r, ceil(k/32)]
    (scale = 2^exponent, value = e4m3(code) * scale)."""
    a = _fortran(a, np.float32)
    r, k = (a.shape[0], a.shape[1]) if rows else (a.shape[1], a.shape[0])
    codes = np.zeros((r, max(k, 1)), np.uint8)
    sexp = np.zeros((r, max((k + 31) // 32, 1)), np.int32)
    lib().orc_quantise_e4m3_codes(r, k, _p(a), max(a.shape[0], 1), _p(codes), _p(sexp), int(rows))
    return codes[:, :k], sexp[:, :(k + 31) // 32]


def gemm_sub(ctype: int, a, b, c, cprec: int = F32) -> np.ndarray:
    """C - A*B with the compute-type semantics of src/gemm.cu."""
    a = _fortran(a, np.float32)
    b = _fortran(b, np.float32)
    c = _fortran(c, np.float32).copy(order="F")
    m, k = a.shape
    n = b.shape[1]
    lib().orc_gemm_sub(ctype, m, n, k, _p(a), max(m, 1), _p(b), max(k, 1), _p(c),
                       max(m, 1), cprec)
    return c


# ---------------------------------------------------------------- drivers
def cfg(W=F32, I=FP32, O=FP32, OP=None, IP=None, OF=None, IF=None, ib=8) -> LuCfg:
    OP = W if OP is None else OP
    IP = OP if IP is None else IP
    OF = OP if OF is None else OF
    IF = OF if IF is None else IF
    return LuCfg(W, I, O, OP, IP, OF, IF, ib)


def lu_nopiv(algo: str, c: LuCfg, a: np.ndarray, nb: int) -> np.ndarray:
    """Run the LL ("ll") or RL ("rl") restatement on a copy of `a`.
    `a` is float32 (W=F32) or uint16 fp16 bits (W=F16)."""
    dt = np.float32 if c.W == F32 else np.uint16
    a = _fortran(a, dt).copy(order="F")
    m, n = a.shape
    f = lib().orc_lu_nopiv_ll_typed if algo == "ll" else lib().orc_lu_nopiv_rl_typed
    st = f(C.byref(c), m, n, nb, _p(a), m)
    if st:
        raise ValueError(f"orc_lu_nopiv_{algo}: bad argument {-st}")
    return a


def lu_f64(a) -> np.ndarray:
    a = _fortran(a, np.float64).copy(order="F")
    st = lib().orc_lu_f64(a.shape[0], _p(a), a.shape[0])
    if st:
        raise ZeroDivisionError(f"zero pivot at column {st}")
    return a


def getrs_f32(lu, b) -> np.ndarray:
    lu = _fortran(lu, np.float32)
    b = _fortran(np.atleast_2d(np.asarray(b, np.float32).T).T, np.float32).copy(order="F")
    lib().orc_getrs_nopiv_f32(lu.shape[0], b.shape[1], _p(lu), lu.shape[0], _p(b), b.shape[0])
    return b


def getrs_f64(lu, b, trans: bool = False) -> np.ndarray:
    lu = _fortran(lu, np.float64)
    b = _fortran(np.atleast_2d(np.asarray(b, np.float64).T).T, np.float64).copy(order="F")
    f = lib().orc_getrs_nopiv_trans_f64 if trans else lib().orc_getrs_nopiv_f64
    f(lu.shape[0], b.shape[1], _p(lu), lu.shape[0], _p(b), b.shape[0])
    return b


# ---------------------------------------------------------------- errors
def _col2d(x, dtype):
    x = np.asarray(x, dtype)
    if x.ndim == 1:
        x = x[:, None]
    return np.asfortranarray(x)


def backward_error(a, b, x) -> float:
    """nwerr, tests/errors.hxx:186-249."""
    a = _fortran(a, np.float32)
    b = _col2d(b, np.float64)
    x = _col2d(x, np.float64)
    m, n = a.shape
    return lib().orc_unsym_backward_error(m, n, _p(a), m, _p(b), b.shape[1], b.shape[0],
                                          _p(x), x.shape[0])


def compwise_error(a, b, x, lu):
    """(cwerr, cwerr_ax), tests/errors.hxx:22-182 (first rhs)."""
    a = _fortran(a, np.float32)
    lu = _fortran(lu, np.float32)
    b = _col2d(b, np.float64)
    x = _col2d(x, np.float64)
    out = np.zeros(1, np.float64)
    m, n = a.shape
    cw = lib().orc_unsym_compwise_error(m, n, _p(a), m, _p(b), _p(x), _p(lu), m, _p(out))
    return cw, float(out[0])


def factor_error(a, lu) -> float:
    """max |A - LU| / (|L||U|), tests/errors.hxx:251-346."""
    a = _fortran(a, np.float32)
    lu = _fortran(lu, np.float32)
    return lib().orc_factor_backward_error(a.shape[0], _p(a), a.shape[0], _p(lu), lu.shape[0])


def residual_f64(a, x, b) -> np.ndarray:
    a = _fortran(a, np.float32)
    x = _col2d(x, np.float64)
    b = _col2d(b, np.float64)
    r = np.zeros_like(b, order="F")
    n = a.shape[0]
    lib().orc_residual_f64(n, b.shape[1], _p(a), n, _p(x), n, _p(b), n, _p(r), n)
    return r


def ir_solve(a, lu, b, tol=1e-14, maxit=30):
    """Iterative refinement with the FP64 residual; returns (x, iters, hist)."""
    a = _fortran(a, np.float32)
    lu = _fortran(lu, np.float32)
    b = _col2d(b, np.float64)
    n, nrhs = b.shape
    x = np.zeros((n, nrhs), np.float64, order="F")
    hist = np.full(maxit + 1, np.nan)
    it = lib().orc_ir_solve(n, nrhs, _p(a), n, _p(lu), n, _p(b), n, _p(x), n, tol,
                            maxit, _p(hist))
    return x, int(it), hist[: (it + 1 if it >= 0 else maxit + 1)]
