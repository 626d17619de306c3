"""CPU tests that pin the oracle (the CPU restatement in oracle/) to facts that
do not come from the oracle itself: the C++ standard library's generator,
numpy/torch casts, exact rational arithmetic, closed forms and invariants.

The reference holds no stored outputs (its tests print, they do not assert),
so these are the checks that make a plausible mistake in the restatement (a
dropped term, a wrong index or sign, a division where the reference
multiplies by a reciprocal, unfused where it fuses) fail."""
import json
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    orc.build()
    orc.set_threads(4)


# ---------------------------------------------------------------- exact rounding
def round_exact(q, p: int, emin: int, emax: int) -> float:
    """Round a rational to the binary format with p significant bits, minimum
    normal exponent emin and maximum exponent emax, to nearest even; overflow
    to +-inf."""
    q = Fraction(q)
    if q == 0:
        return 0.0
    s = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    quantum = Fraction(2) ** (max(e, emin) - (p - 1))
    n, rem = divmod(a, quantum)
    n = int(n)
    if rem * 2 > quantum or (rem * 2 == quantum and n % 2 == 1):
        n += 1
    r = n * quantum
    if r >= Fraction(2) ** (emax + 1):
        return s * float("inf")
    return s * float(r)


def rf32(q):
    return round_exact(q, 24, -126, 127)


def rf16(q):
    return round_exact(q, 11, -14, 15)


# ---------------------------------------------------------------- generator
def _hex_to_f32(hs):
    return np.array([int(h, 16) for h in hs], np.uint32).view(np.float32)


def test_minstd_matches_libstdcxx():
    """orc_minstd_stream reproduces std::default_random_engine +
    uniform_real_distribution<float>(-1, 1) bit for bit (tests/genmat.hxx:72-73)."""
    gold = json.load(open(os.path.join(GOLDEN, "minstd.json")))
    for seed, hs in gold.items():
        want = _hex_to_f32(hs)
        got = orc.minstd_stream(1 if int(seed) < 0 else int(seed), len(want))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), seed


def test_gen_diagdom_layout():
    """Draws fill A row by row (i outer, j inner) and a_ii = m
    (tests/genmat.hxx:86-97)."""
    m, n = 7, 5
    a = orc.gen_diagdom(m, n)
    s = orc.minstd_stream(1, m * n).reshape(m, n)
    for i in range(m):
        for j in range(n):
            want = float(m) if i == j else s[i, j]
            assert a[i, j] == want


def test_gen_diagdom_alpha_diagonal():
    n, seed, alpha = 64, 2024, 1.05
    a = orc.gen_diagdom_alpha(n, seed, alpha)
    s = orc.minstd_stream(seed, n * n).reshape(n, n).astype(np.float64)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(a[off], s.astype(np.float32)[off])
    rows = np.where(off, np.abs(s), 0.0).sum(axis=1)  # exact: dyadic values
    want = np.float32(alpha * rows)
    assert np.array_equal(np.diag(a), want)
    # row diagonal dominance by the factor alpha
    assert np.all(np.abs(np.diag(a)) >= alpha * rows * (1 - 2 ** -23))


def test_minstd_jump_ahead_consistent():
    """the per-row jump-ahead the parallel generator uses equals sequential draws"""
    n = 33
    a = orc.gen_diagdom_alpha(n, 99, 2.0)
    s = orc.minstd_stream(99, n * n).reshape(n, n)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(a[off], s[off])


def test_gen_uniform_row_major_draws():
    """right-hand sides are drawn like gen_diagdom: row by row"""
    m, n = 6, 3
    b = orc.gen_uniform(m, n, 5)
    s = orc.minstd_stream(5, m * n)
    assert np.array_equal(b.flatten(order="C"), s)


# ---------------------------------------------------------------- conversions
def test_f16_widening_exhaustive():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = orc.f16_bits_to_f32(h.reshape(256, 256))
    want = h.view(np.float16).astype(np.float32).reshape(256, 256)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_narrowing_truth_table():
    g = json.load(open(os.path.join(GOLDEN, "convert.json")))
    x = np.array(g["f32_bits"], np.uint32).view(np.float32)[None, :]
    assert np.array_equal(orc.convert(x, orc.F32, orc.F16)[0], np.array(g["f16_bits"], np.uint16))
    assert np.array_equal(orc.convert(x, orc.F32, orc.BF16)[0],
                          np.array(g["bf16_bits"], np.uint16))
    xe = np.array(g["e4m3_f32_bits"], np.uint32).view(np.float32)[None, :]
    assert np.array_equal(orc.convert(xe, orc.F32, orc.E4M3)[0], np.array(g["e4m3_bits"], np.uint8))


def test_f32_to_f16_random_patterns_vs_numpy():
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, size=(512, 512), dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        want = x.astype(np.float16)
    got = orc.convert(x, orc.F32, orc.F16).view(np.float16)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint16), want[~nan].view(np.uint16))


def test_f64_to_f16_direct_rounding():
    """double -> half rounds once (numpy's direct cast), not via float."""
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((64, 64)) * 10.0 ** rng.integers(-9, 5, (64, 64)))
    # a value whose double->float->half double rounding differs
    x[0, 0] = 1.0 + 2.0 ** -11 + 2.0 ** -40
    want = x.astype(np.float16)
    got = orc.convert(x, orc.F64, orc.F16).view(np.float16)
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))
    assert got[0, 0] == np.float16(1.0 + 2.0 ** -10)


def test_e4m3_decode_all_codes():
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    want = torch.from_numpy(codes.copy()).view(torch.float8_e4m3fn).float().numpy()
    got = orc.convert(codes[None, :], orc.E4M3, orc.F32)[0]
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan], want[~nan])


def test_e4m3_saturates():
    x = np.array([[1000.0, -1000.0, 449.0, np.inf]], np.float32)
    got = orc.convert(orc.convert(x, orc.F32, orc.E4M3), orc.E4M3, orc.F32)[0]
    assert list(got[:3]) == [448.0, -448.0, 448.0]


# ---------------------------------------------------------------- panel
def test_panel_rcp_then_multiply_fused_update():
    """2x2 tile: l21 = a21 * rn(1/a11) (not a21/a11) and u22 = fma(-l21, a12, a22)
    rounded once (src/lu_nopiv.cu:138-167)."""
    a11, a12, a21 = np.float32(3.0), np.float32(1.0 + 2.0 ** -20), np.float32(5.0)
    a22 = np.float32(1.0)
    d = np.array([[a11, a12], [a21, a22]], np.float32)
    l, _, stat = orc.lu_nopiv_panel(orc.F32, 8, 2, 2, d, d.copy(), np.zeros((2, 2), np.float32))
    r = rf32(Fraction(1) / Fraction(float(a11)))
    l21 = rf32(Fraction(float(a21)) * Fraction(r))
    assert l21 != rf32(Fraction(5, 3))  # the two orders differ here
    assert l[1, 0] == np.float32(l21)
    u22 = rf32(Fraction(float(a22)) - Fraction(l21) * Fraction(float(a12)))
    unfused = np.float32(np.float32(a22) - np.float32(np.float32(l21) * a12))
    assert u22 != float(unfused)
    assert l[1, 1] == np.float32(u22)
    assert stat == 2


def test_panel_f16_rounds_each_operation():
    a11, a21 = 3.0, 5.0
    d = np.array([[a11, 1.0], [a21, 1.0]])
    dh = orc.f32_to_f16_bits(d.astype(np.float32))
    l, _, _ = orc.lu_nopiv_panel(orc.F16, 8, 2, 2, dh, dh.copy(), np.zeros((2, 2), np.uint16))
    r = rf16(Fraction(1, 3))
    want = rf16(Fraction(5) * Fraction(r))
    assert orc.f16_bits_to_f32(l)[1, 0] == np.float32(want)


def _exact_factors(n, seed, ib_rows=0):
    """unit lower L0 and upper U0 with dyadic entries (|l| <= 1/2 on a band,
    u_kk a power of two) so every product and partial sum of the elimination
    is exact in fp16 and fp32."""
    rng = np.random.default_rng(seed)
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - 2), i):
            L[i, j] = rng.choice([-0.5, 0.0, 0.25, 0.5])
        U[i, i] = rng.choice([2.0, 4.0, -2.0])
        for j in range(i + 1, min(n, i + 3)):
            U[i, j] = rng.choice([-1.0, 0.0, 0.5, 1.0])
    return L, U


@pytest.mark.parametrize("prec", [orc.F32, orc.F16])
def test_panel_exact_arithmetic(prec):
    ib, m = 8, 24
    L, U = _exact_factors(m, 3)
    A = L @ U
    d = A[:ib, :ib]
    if prec == orc.F16:
        cv = lambda x: orc.f32_to_f16_bits(x.astype(np.float32))
        back = orc.f16_bits_to_f32
    else:
        cv = lambda x: x.astype(np.float32)
        back = lambda x: x
    l_in = A[:, :ib]
    u_in = A[:ib, :]
    l, u, _ = orc.lu_nopiv_panel(prec, ib, m, ib, cv(d), cv(l_in), cv(u_in))
    lu_tile = np.tril(L[:ib, :ib], -1) + np.triu(U[:ib, :ib])
    assert np.array_equal(back(l)[:ib], lu_tile.astype(np.float32))
    assert np.array_equal(back(l)[ib:], L[ib:, :ib].astype(np.float32))
    # U12 = L11^-1 A12 for the columns right of the tile
    assert np.array_equal(back(u)[:, ib:], U[:ib, ib:].astype(np.float32))


def test_panel_matches_f64_doolittle():
    rng = np.random.default_rng(4)
    ib, m = 16, 200
    A = rng.uniform(-1, 1, (m, m)).astype(np.float32)
    A[np.arange(ib), np.arange(ib)] += 8.0
    l, u, _ = orc.lu_nopiv_panel(orc.F32, ib, m, ib, A[:ib, :ib], A[:, :ib], A[:ib, :])
    # textbook Doolittle of the first ib columns in double
    T = A[:ib, :ib].astype(np.float64)
    Lt = np.eye(ib)
    Ut = np.zeros((ib, ib))
    for k in range(ib):
        Ut[k, k:] = T[k, k:] - Lt[k, :k] @ Ut[:k, k:]
        Lt[k + 1:, k] = (T[k + 1:, k] - Lt[k + 1:, :k] @ Ut[:k, k]) / Ut[k, k]
    L21 = A[ib:, :ib].astype(np.float64) @ np.linalg.inv(Ut)
    U12 = np.linalg.solve(Lt, A[:ib, ib:].astype(np.float64))
    tol = 64 * 2.0 ** -24
    assert np.allclose(l[:ib], np.tril(Lt, -1) + Ut, rtol=0, atol=tol * 8)
    assert np.allclose(l[ib:], L21, rtol=0, atol=tol * 8)
    assert np.allclose(u[:, ib:], U12, rtol=0, atol=tol * 8)


# ---------------------------------------------------------------- GEMM
def test_gemm_exact_small_integers():
    rng = np.random.default_rng(6)
    a = rng.integers(-4, 5, (37, 29)).astype(np.float32)
    b = rng.integers(-4, 5, (29, 23)).astype(np.float32)
    c = rng.integers(-50, 50, (37, 23)).astype(np.float32)
    want = c - a.astype(np.float64) @ b
    for ct in (orc.FP32, orc.TC32, orc.FP16, orc.TC16, orc.BF16TC32):
        assert np.array_equal(orc.gemm_sub(ct, a, b, c), want.astype(np.float32)), ct


@pytest.mark.parametrize("ct,opp", [(orc.FP32, orc.F32), (orc.TC32, orc.F16),
                                    (orc.BF16TC32, orc.BF16)])
def test_gemm_error_bound(ct, opp):
    """|C - C^| <= gamma_{k+1} (|A~||B~| + |C|) with the operands rounded to the
    compute type's input precision (A~, B~) and fp32 accumulation."""
    rng = np.random.default_rng(7)
    m, n, k = 64, 48, 300
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    at = orc.round_to(opp, a).astype(np.float64)
    bt = orc.round_to(opp, b).astype(np.float64)
    exact = c - at @ bt
    got = orc.gemm_sub(ct, a, b, c)
    u = 2.0 ** -24
    g = (k + 1) * u / (1 - (k + 1) * u)
    bound = g * (np.abs(at) @ np.abs(bt) + np.abs(c))
    assert np.all(np.abs(got - exact) <= bound)
    # and the operand rounding is really applied: unrounded operands miss
    if opp != orc.F32:
        near = np.max(np.abs(got - exact))
        assert np.max(np.abs(got - (c - a.astype(np.float64) @ b))) > 100 * near


def test_gemm_fp16_accumulation_brute_force():
    """FP16 compute: products summed from zero in k order, each step one fp16
    rounding, then C - acc rounded to C's precision."""
    rng = np.random.default_rng(8)
    m, n, k = 3, 2, 40
    a = orc.round_to(orc.F16, rng.uniform(-1, 1, (m, k)))
    b = orc.round_to(orc.F16, rng.uniform(-1, 1, (k, n)))
    c = orc.round_to(orc.F16, rng.uniform(-1, 1, (m, n)))
    got = orc.gemm_sub(orc.FP16, a, b, c, orc.F16)
    for i in range(m):
        for j in range(n):
            acc = 0.0
            for q in range(k):
                acc = rf16(Fraction(float(a[i, q])) * Fraction(float(b[q, j])) + Fraction(acc))
            want = rf16(Fraction(float(c[i, j])) - Fraction(acc))
            assert got[i, j] == np.float32(want)


def test_quantise_e4m3_known_answers():
    """Hand-derived codes of the FP8TC32 operand format: scale 2^(floor(log2
    amax) - 8) per 32-block, e4m3 nearest even, saturation at 448."""
    row = np.zeros(70, np.float32)
    row[:6] = [1.0, 0.75, 2.0**-17, 3 * 2.0**-19, -0.5, 5 * 2.0**-19]
    row[32:34] = [1.9, -1.0]   # block 1: amax 1.9 -> 1.9 * 2^8 = 486.4 saturates
    row[64:66] = [3.0, 0.25]   # block 2 (partial, 6 long)
    codes, sexp = orc.quantise_e4m3(row[None, :], rows=True)
    assert sexp.tolist() == [[-8, -8, -7]]
    # 1.0 * 2^8 = 256 = 1.0 * 2^8 -> e 15 m 0; 0.75 * 2^8 = 192 = 1.5 * 2^7 -> e 14 m 4
    # 2^-17 * 2^8 = 2^-9 -> subnormal m 1; 3 * 2^-11 = 0.75 * 2^-9 -> rounds to 2^-9
    # -0.5 * 2^8 = -128 -> sign, e 14 m 0; 5 * 2^-11 = 1.25 * 2^-9 -> 2^-9 (nearest)
    assert codes[0, :6].tolist() == [0x78, 0x74, 0x01, 0x01, 0xF0, 0x01]
    # 486.4 -> 448 (0x7E); -256 -> 0xF8
    assert codes[0, 32:34].tolist() == [0x7E, 0xF8]
    # block 2: amax 3 -> 2^(1 - 8): 3 * 2^7 = 384 = 1.5 * 2^8 -> e 15 m 4; 0.25 * 2^7 = 32 -> e 12
    assert codes[0, 64:66].tolist() == [0x7C, 0x60]
    assert np.all(codes[0, 6:32] == 0) and np.all(codes[0, 66:] == 0)
    # an all-zero block has scale 2^0
    z = orc.quantise_e4m3(np.zeros((1, 40), np.float32))
    assert z[1].tolist() == [[0, 0]] and not z[0].any()
    # rows of A and columns of the transpose are the same blocks
    rng = np.random.default_rng(4)
    a = (rng.standard_normal((7, 100)) * np.exp2(rng.integers(-20, 10, (7, 1)))).astype(np.float32)
    ca, ea = orc.quantise_e4m3(a, rows=True)
    cb, eb = orc.quantise_e4m3(np.asfortranarray(a.T), rows=False)
    assert np.array_equal(ca, cb) and np.array_equal(ea, eb)


def test_quantise_e4m3_elements_match_torch_float8():
    """Element rounding against torch's float8_e4m3fn cast (nearest even)
    where the scaled value is within the finite range (torch does not
    saturate: values above 448 are the oracle's saturation case)."""
    import torch
    rng = np.random.default_rng(12)
    a = (rng.standard_normal((64, 256)) * np.exp2(rng.integers(-12, 12, (64, 8)).repeat(32, 1))).astype(np.float32)
    codes, sexp = orc.quantise_e4m3(a, rows=True)
    scaled = a * np.exp2(-sexp.repeat(32, axis=1).astype(np.float64))
    want = torch.from_numpy(scaled.astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    inr = np.abs(scaled) <= 448
    assert np.array_equal(codes[inr], want[inr])
    assert np.all((codes[~inr] & 0x7F) == 0x7E)
    assert inr.mean() > 0.9


def test_gemm_fp8_per_block_scale_bound():
    rng = np.random.default_rng(9)
    m, n, k = 32, 32, 128
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    c = np.zeros((m, n), np.float32)
    got = orc.gemm_sub(orc.FP8TC32, a, b, c)
    exact = -(a.astype(np.float64) @ b)
    # e4m3 keeps 4 significant bits: relative operand error <= 2^-4
    bound = 2 * 2.0 ** -4 * (np.abs(a) @ np.abs(b)) + 1e-5
    assert np.all(np.abs(got - exact) <= bound)


# ---------------------------------------------------------------- drivers
CFGS = [
    ("f32", dict(W=orc.F32, I=orc.FP32, O=orc.FP32)),
    ("f32_tc32", dict(W=orc.F32, I=orc.TC32, O=orc.TC32)),
    ("f16_tc32_fp32buf", dict(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32)),
    ("f16_tc32", dict(W=orc.F16, I=orc.TC32, O=orc.TC32)),
    ("f16_fp16", dict(W=orc.F16, I=orc.FP16, O=orc.FP16)),
]


def _prep(kw, a):
    return orc.convert(a.astype(np.float32), orc.F32, orc.F16) if kw["W"] == orc.F16 else a


def _unprep(kw, a):
    return orc.f16_bits_to_f32(a) if kw["W"] == orc.F16 else a


@pytest.mark.parametrize("name,kw", CFGS)
@pytest.mark.parametrize("n,nb,ib", [(24, 16, 8), (100, 32, 8), (64, 64, 16)])
def test_drivers_exact_arithmetic(name, kw, n, nb, ib):
    """A = L0 U0 with every intermediate exact: LL and RL return L0 \\ U0."""
    L, U = _exact_factors(n, n)
    A = (L @ U).astype(np.float32)
    want = (np.tril(L, -1) + U).astype(np.float32)
    for algo in ("ll", "rl"):
        if algo == "ll" or kw.get("OP", kw["W"]) == kw["W"]:
            c = orc.cfg(ib=ib, **kw)
            got = _unprep(kw, orc.lu_nopiv(algo, c, _prep(kw, A), nb))
            assert np.array_equal(got, want), (algo, name)


def _lu_parts(lu):
    return np.tril(lu, -1) + np.eye(lu.shape[0]), np.triu(lu)


def test_ll_f32_componentwise_bound_and_growth():
    n, nb = 256, 64
    a = orc.gen_diagdom_alpha(n, 3, 1.5)
    lu = orc.lu_nopiv("ll", orc.cfg(ib=8), a, nb)
    L, U = _lu_parts(lu.astype(np.float64))
    u = 2.0 ** -24
    g = n * u / (1 - n * u)
    assert np.all(np.abs(a - L @ U) <= g * (np.abs(L) @ np.abs(U)))
    # growth factor of a diagonally dominant matrix is at most 2
    assert np.max(np.abs(U)) <= 2 * np.max(np.abs(a))


def test_ll_rl_agree_with_f64_lu():
    n, nb = 192, 64
    a = orc.gen_diagdom_alpha(n, 12, 1.1)
    ref = orc.lu_f64(a)
    for algo in ("ll", "rl"):
        lu = orc.lu_nopiv(algo, orc.cfg(ib=8), a, nb)
        err = np.max(np.abs(lu - ref)) / np.max(np.abs(ref))
        assert err < 64 * 2.0 ** -24, algo


def test_f64_lu_textbook():
    rng = np.random.default_rng(10)
    a = rng.uniform(-1, 1, (9, 9)) + 9 * np.eye(9)
    lu = orc.lu_f64(a)
    L, U = _lu_parts(lu)
    assert np.allclose(L @ U, a, atol=1e-14)


def test_fp16_storage_error_scales_with_fp16_unit_roundoff():
    n, nb = 256, 64
    a = orc.gen_diagdom_alpha(n, 5, 1.05)
    a16 = orc.convert(a, orc.F32, orc.F16)
    ref = orc.lu_f64(orc.f16_bits_to_f32(a16))
    c = orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32)
    lu = orc.f16_bits_to_f32(orc.lu_nopiv("ll", c, a16, nb))
    u16 = 2.0 ** -11
    for part in (lambda x: np.tril(x, -1), np.triu):
        err = np.max(np.abs(part(lu - ref))) / np.max(np.abs(part(ref)))
        # stored in fp16: about one fp16 rounding, far above fp32's
        assert u16 / 8 < err < 4 * u16
    lu32 = orc.lu_nopiv("ll", orc.cfg(ib=8), orc.f16_bits_to_f32(a16), nb)
    assert np.max(np.abs(lu32 - ref)) / np.max(np.abs(ref)) < 64 * 2.0 ** -24


def test_drivers_bad_arguments():
    a = np.eye(8, dtype=np.float32)
    with pytest.raises(ValueError):
        orc.lu_nopiv("ll", orc.cfg(ib=0), a, 4)


# ---------------------------------------------------------------- errors
def test_backward_error_closed_form_identity():
    """A = I, b = x: nwerr = ||x^ - x||_inf / (||x^||_inf + ||x||_inf) (tests/errors.hxx:237)."""
    n = 5
    x = np.array([1.0, -2.0, 0.5, 3.0, -1.0])
    e = np.array([1e-3, 0, -2e-3, 0, 0])
    xh = x + e
    got = orc.backward_error(np.eye(n, dtype=np.float32), x, xh)
    want = np.max(np.abs(e)) / (np.max(np.abs(xh)) + np.max(np.abs(x)))
    assert got == pytest.approx(want, rel=1e-14)
    assert orc.backward_error(np.eye(n, dtype=np.float32), x, x) == 0.0


def test_backward_error_scale_invariant():
    rng = np.random.default_rng(11)
    a = rng.uniform(-1, 1, (20, 20)).astype(np.float32) + 20 * np.eye(20, dtype=np.float32)
    x = rng.uniform(-1, 1, 20)
    b = a.astype(np.float64) @ np.ones(20)
    e1 = orc.backward_error(a, b, x)
    e2 = orc.backward_error(4 * a, 4 * b, x)
    assert e1 == pytest.approx(e2, rel=1e-12)


def test_compwise_error_closed_form_identity():
    """A = LU = I: cwerr = max_i |e_i| / (2|x^_i|), Oettli-Prager form
    max_i |e_i| / |x^_i| (tests/errors.hxx:147-166)."""
    x = np.array([1.0, -2.0, 0.5])
    e = np.array([1e-3, -1e-3, 1e-4])
    xh = x + e
    eye = np.eye(3, dtype=np.float32)
    cw, cw_ax = orc.compwise_error(eye, x, xh, eye)
    r = np.abs(e) / np.abs(xh)
    assert cw == pytest.approx(np.max(r) / 2, rel=1e-12)
    assert cw_ax == pytest.approx(np.max(r), rel=1e-12)


def test_factor_error_zero_for_exact_factors():
    L, U = _exact_factors(12, 1)
    a = (L @ U).astype(np.float32)
    lu = (np.tril(L, -1) + U).astype(np.float32)
    assert orc.factor_error(a, lu) == 0.0
    lu2 = lu.copy()
    lu2[5, 3] += 0.25
    assert orc.factor_error(a, lu2) > 0.0


def test_factor_error_hand_computed():
    """max_ij |A - LU|_ij / (|L||U|)_ij on cases worked by hand
    (tests/errors.hxx:251-346). 2x2: A = [[4, 2], [2, 3]] with l21 = 3/4,
    u22 = 9/4: |A - LU| = [[0, 0], [1, 3/4]], |L||U| = [[4, 2], [3, 15/4]], so
    the error is 1/3 (dividing by |A| would give 1/2). 3x3 with mixed signs,
    L = [[1,0,0],[-1/2,1,0],[1/4,1,1]], U = [[2,-1,1],[0,3,-2],[0,0,4]]: LU =
    [[2,-1,1],[-1,7/2,-5/2],[1/2,11/4,9/4]], |L||U| = [[2,1,1],[1,7/2,5/2],
    [1/2,13/4,25/4]]; A = LU except a_21 += 1/2 and a_12 -= 1/4, so the
    error is max(1/2 / 13/4, 1/4 / 5/2) = 2/13 (|LU| in the denominator would
    give 2/11)."""
    a = np.array([[4, 2], [2, 3]], np.float32)
    lu = np.array([[4, 2], [0.75, 2.25]], np.float32)
    assert orc.factor_error(a, lu) == pytest.approx(1.0 / 3.0, rel=1e-15)
    lu3 = np.array([[2, -1, 1], [-0.5, 3, -2], [0.25, 1, 4]], np.float32)
    a3 = np.array([[2, -1, 1], [-1, 3.5, -2.75], [0.5, 3.25, 2.25]], np.float32)
    assert orc.factor_error(a3, lu3) == pytest.approx(2.0 / 13.0, rel=1e-15)


# ---------------------------------------------------------------- parity rule
def _ll16(a16, nb, **kw):
    c = dict(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=32)
    c.update(kw)
    return orc.lu_nopiv("ll", orc.cfg(**c), a16, nb)


@pytest.mark.parametrize("n,nb,alpha", [(256, 64, 1.5), (700, 128, 1.1)])
def test_parity_rule_rejects_planted_rounding_mistakes(n, nb, alpha):
    """The entrywise rule the GPU parity tests apply (tests/helpers.py)
    accepts two oracle results that differ only in summation order (k
    ascending vs descending: what the matrix cores change) and rejects
    factors computed with one rounding point moved, planted here in the
    oracle itself: fp16 outer panel buffers where the headline layout keeps
    fp32 ones, bf16 instead of fp16 outer operands, unrounded (fp32) instead
    of fp16 inner operands; for fp32 storage, fp16-rounded outer operands
    where fp32 ones are used and the reverse."""
    from helpers import parity_ok
    a32 = orc.gen_diagdom_alpha(n, 11, alpha)
    a16 = orc.f32_to_f16_bits(a32)
    ref = _ll16(a16, nb)
    f32 = lambda I, O: orc.lu_nopiv("ll", orc.cfg(W=orc.F32, I=I, O=O, ib=8), a32, nb)
    r32 = f32(orc.FP32, orc.TC32)
    try:
        orc.set_sum_order(1)
        rev16, rev32 = _ll16(a16, nb), f32(orc.FP32, orc.TC32)
    finally:
        orc.set_sum_order(0)
    assert not np.array_equal(rev32, r32)  # the order does change the results
    for g, o in ((ref, ref), (rev16, ref), (r32, r32), (rev32, r32)):
        ok, st = parity_ok(g, o)
        assert ok, st
    for wrong in (dict(OP=orc.F16, IP=orc.F32, OF=orc.F16, IF=orc.F32), dict(O=orc.BF16TC32),
                  dict(I=orc.FP32)):
        ok, st = parity_ok(_ll16(a16, nb, **wrong), ref)
        assert not ok, (wrong, st)
    for I, O in ((orc.FP32, orc.FP32), (orc.TC32, orc.TC32)):
        ok, st = parity_ok(f32(I, O), r32)
        assert not ok, (I, O, st)


def test_entry_scale_closed_form():
    """s = |L||U| (divided by |u_jj| below the diagonal) on a 2x2 case:
    L = [[1, 0], [-1/2, 1]], U = [[4, -2], [0, 3]]: |L||U| = [[4, 2], [2, 4]],
    so s = [[4, 2], [2 / 4, 4]]"""
    from helpers import entry_scale
    lu = np.array([[4, -2], [-0.5, 3]], np.float32)
    assert np.allclose(entry_scale(lu), [[4, 2], [0.5, 4]], rtol=0, atol=0)


def test_ulp_distance_known_values():
    from helpers import ulp_distance
    h = orc.f32_to_f16_bits(np.array([1.0, 1.0, -0.0, 65504.0, -1.0], np.float32))
    h2 = orc.f32_to_f16_bits(np.array([1.0 + 2**-10, 1.0 - 2**-11, 0.0, 32768.0, 1.0],
                                      np.float32))
    assert ulp_distance(h, h2).ravel().tolist() == [1, 1, 0, 0x7BFF - 0x7800, 2 * 0x3C00]
    f = np.array([1.0, -0.0, 2.0], np.float32)
    f2 = np.array([np.nextafter(np.float32(1), np.float32(2)), 0.0, 2.0], np.float32)
    assert ulp_distance(f, f2).ravel().tolist() == [1, 0, 0]


# ---------------------------------------------------------------- residual / IR
def test_residual_f64_vs_exact():
    rng = np.random.default_rng(12)
    n = 24
    a = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    x = rng.uniform(-1, 1, n)
    b = rng.uniform(-1, 1, n)
    r = orc.residual_f64(a, x, b)[:, 0]
    for i in range(n):
        exact = Fraction(float(b[i])) - sum(Fraction(float(a[i, j])) * Fraction(float(x[j]))
                                            for j in range(n))
        scale = abs(b[i]) + np.abs(a[i].astype(np.float64)) @ np.abs(x)
        assert abs(r[i] - float(exact)) <= (n + 1) * 2.0 ** -53 * scale


def test_getrs_f64_and_transposed():
    rng = np.random.default_rng(13)
    n = 30
    a = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    lu = orc.lu_f64(a)
    b = rng.uniform(-1, 1, (n, 3))
    x = orc.getrs_f64(lu, b)
    assert np.allclose(a @ x, b, atol=1e-13)
    xt = orc.getrs_f64(lu, b, trans=True)
    assert np.allclose(a.T @ xt, b, atol=1e-13)


def test_ir_trace_fixture():
    g = json.load(open(os.path.join(GOLDEN, "ir_trace.json")))
    n, seed = g["n"], g["seed"]
    a = orc.gen_diagdom_alpha(n, seed, g["alpha"])
    a16 = orc.convert(a, orc.F32, orc.F16)
    c = orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=g["ib"])
    lu = orc.f16_bits_to_f32(orc.lu_nopiv("ll", c, a16, g["nb"]))
    b = orc.gen_uniform(n, 1, seed + 1).astype(np.float64)
    x, it, hist = orc.ir_solve(a, lu, b, tol=1e-14, maxit=30)
    assert it == g["iters"]
    assert np.allclose(hist, g["history"], rtol=1e-6, atol=0)
    # converged: the restated error measure agrees
    assert orc.backward_error(a, b, x) <= 1e-14
    # refinement contracts by about the fp16 factorization's error each step
    assert all(h1 < h0 * 1e-2 for h0, h1 in zip(hist[:-1], hist[1:]))


# ---------------------------------------------------------------- generators / equilibration
def _fnv1a(a: np.ndarray) -> str:
    h = 1469598103934665603
    for b in np.asfortranarray(a, np.float32).ravel(order="F").view(np.uint8):
        h ^= int(b)
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def test_arrowhead_and_randmat_match_libstdcxx():
    """gen_arrowhead / gen_randmat (tests/common.hxx:298-319, 391-455) of the
    oracle against the same loops run with libstdc++'s engine and
    uniform_real_distribution<float> (tests/golden/minstd_ref.cpp)"""
    g = json.load(open(os.path.join(GOLDEN, "generators.json")))
    for n, lo, hi, h in g["arrowhead"]:
        assert _fnv1a(orc.gen_arrowhead(n, lo, hi)) == h, (n, lo, hi)
    for m, n, seed, h in g["randmat"]:
        assert _fnv1a(orc.gen_randmat(m, n, seed)) == h, (m, n, seed)


def test_arrowhead_structure():
    a = orc.gen_arrowhead(600, -1.0, 1.0)
    assert np.all(np.diag(a) == 512.0)
    nz = a != 0
    assert not nz[0:256, 256:344].any() and not nz[256:344, 0:256].any()  # off-block zeros
    assert nz[344:, :].mean() > 0.99 and nz[:, 344:].mean() > 0.99      # the arrow
    assert np.all(np.abs(a[a != 512.0]) < 1.0)


def test_geequb_matches_lapack():
    """the oracle's restatement of sgeequb against scipy's bundled LAPACK on
    badly scaled matrices (tests/golden/geequb.json): bit-identical scale
    factors, same rowcnd / colcnd / amax / info"""
    g = json.load(open(os.path.join(GOLDEN, "geequb.json")))
    for c in g["cases"]:
        m, n = c["m"], c["n"]
        a = np.array(c["a_bits"], np.uint32).view(np.float32).reshape((m, n), order="F")
        r, cc, rowcnd, colcnd, amax, info = orc.geequb(a)
        assert info == c["info"]
        assert r.view(np.uint32).tolist() == c["r_bits"]
        assert cc.view(np.uint32).tolist() == c["c_bits"]
        assert (rowcnd, colcnd, amax) == (np.float32(c["rowcnd"]), np.float32(c["colcnd"]),
                                          np.float32(c["amax"]))


def test_geequb_zero_row_and_column_info():
    a = np.asfortranarray(np.eye(4, dtype=np.float32))
    a[2, 2] = 0.0
    assert orc.geequb(a)[5] == 3
    b = np.asfortranarray(np.ones((3, 3), np.float32))
    b[:, 1] = 0.0
    assert orc.geequb(b)[5] == 3 + 2


def test_laqge_branches_closed_form():
    """slaqge: no scaling when rowcnd and colcnd >= 0.1 (and amax in range),
    columns only when just colcnd < 0.1, rows only when just rowcnd < 0.1,
    both otherwise; power-of-two factors multiply exactly"""
    a = np.asfortranarray(np.array([[3.0, 5.0], [7.0, 11.0]], np.float32))
    r = np.array([0.5, 0.25], np.float32)
    c = np.array([2.0, 0.125], np.float32)
    out, eq = orc.laqge(a, r, c, 1.0, 1.0, 11.0)
    assert eq == "N" and np.array_equal(out, a)
    out, eq = orc.laqge(a, r, c, 1.0, 0.05, 11.0)
    assert eq == "C" and np.array_equal(out, a * c[None, :])
    out, eq = orc.laqge(a, r, c, 0.05, 1.0, 11.0)
    assert eq == "R" and np.array_equal(out, a * r[:, None])
    out, eq = orc.laqge(a, r, c, 0.05, 0.05, 11.0)
    assert eq == "B" and np.array_equal(out, a * r[:, None] * c[None, :])
    out, eq = orc.laqge(a, r, c, 1.0, 1.0, 2.0**-110)  # amax below SMALL: rows scaled
    assert eq == "R"


def test_read_mtx_symmetric_fixture():
    """the harness's MatrixMarket reader (tests/genmat.hxx:16-63): every
    stored entry lands at (row, col) and (col, row)"""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from testing_lu_gpu import read_mtx
    a, nrows = read_mtx(os.path.join(GOLDEN, "small_sym.mtx"))
    want = np.array([[4, -1.5, 0, 0, 2], [-1.5, 5.25, 0, 0.125, 0], [0, 0, 6, 0, 0],
                     [0, 0.125, 0, 7, 0], [2, 0, 0, 0, 8.5]], np.float32)
    assert nrows == 5 and np.array_equal(a, want)
