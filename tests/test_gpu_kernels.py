"""GPU parity of the kernels behind the C ABI against the CPU oracle.

Bit-exact: the generator, every conversion, the panel kernel (same operation
order as src/lu_nopiv.cu) and the fp16-accumulating GEMM (one fp16 fma per k
in k order). Matrix-core GEMMs sum in a different order than the oracle, so
they are checked against the error bound 2*gamma_k*|A||B| (+ output rounding)
that both results satisfy with respect to the exact product.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import to_dev, to_np

pytestmark = pytest.mark.gpu


def _rf():
    import remifa_amd as rf
    return rf


# ---------------------------------------------------------------- generator
@pytest.mark.parametrize("n,seed,alpha", [(1, 3, 1.5), (7, 1, 0.0), (256, 42, 1.5), (300, 7, 1.01),
                                          (1031, 2, 1.05), (256, 5, 0.0)])
def test_gen_diagdom_bitexact(handle, gpu, n, seed, alpha):
    import torch
    rf = _rf()
    a = rf.colmajor_empty(n, n, device=gpu, ld=n + 3)
    rf.gen_diagdom(handle, a, seed=seed, alpha=alpha)
    torch.cuda.synchronize()
    ref = orc.gen_diagdom(n) if alpha <= 0 else orc.gen_diagdom_alpha(n, seed, alpha)
    assert np.array_equal(to_np(a).view(np.uint32), ref.view(np.uint32))


def test_gen_uniform_bitexact(handle, gpu):
    import torch
    rf = _rf()
    b = rf.colmajor_empty(300, 16, device=gpu)
    rf.gen_uniform(handle, b, seed=99)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(b).view(np.uint32), orc.gen_uniform(300, 16, 99).view(np.uint32))


# ---------------------------------------------------------------- convert
def _special_f32(rng, n):
    bits = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x)]
    extra = np.array([0.0, -0.0, 1.0, -1.0, 65504, 65519.99, 65520, 65536, 2**-24, 2**-25,
                      3 * 2**-26, 2**-14, 448, 464, 480, 2**-9, 2**-10, np.inf, -np.inf,
                      1 + 2**-11, 1 + 3 * 2**-11, 3.4e38], np.float32)
    return np.concatenate([x, extra, (rng.standard_normal(4000) * 50).astype(np.float32)])


@pytest.mark.parametrize("dst", ["f16", "bf16", "e4m3"])
def test_convert_f32_narrowing_bitexact(handle, gpu, dst):
    import torch
    rf = _rf()
    rng = np.random.default_rng(0)
    x = _special_f32(rng, 20000)
    m = 97
    x = x[: (len(x) // m) * m].reshape(m, -1, order="F")
    code = {"f16": (rf.R_16F, orc.F16, torch.float16), "bf16": (rf.R_16BF, orc.BF16, torch.bfloat16),
            "e4m3": (rf.R_8F_E4M3, orc.E4M3, torch.float8_e4m3fn)}[dst]
    a = to_dev(x, gpu, ld=m + 5)
    out = rf.colmajor_empty(m, x.shape[1], dtype=code[2], device=gpu, ld=m + 1)
    rf.convert(handle, a, out)
    torch.cuda.synchronize()
    got = to_np(out)
    ref = orc.convert(x, orc.F32, code[1])
    assert np.array_equal(got, ref)


def test_convert_widening_exhaustive(handle, gpu):
    import torch
    rf = _rf()
    h = np.arange(65536, dtype=np.uint16).reshape(256, 256, order="F")
    a = to_dev(h, gpu)
    out = rf.colmajor_empty(256, 256, device=gpu)
    rf.convert(handle, a, out)
    torch.cuda.synchronize()
    got = to_np(out)
    ref = orc.convert(h, orc.F16, orc.F32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    # f16 -> f64 and back to f16 is the identity on non-NaN patterns
    out64 = rf.colmajor_empty(256, 256, dtype=torch.float64, device=gpu)
    rf.convert(handle, a, out64)
    back = rf.colmajor_empty(256, 256, dtype=torch.float16, device=gpu)
    rf.convert(handle, out64, back)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(back)[~nan], h[~nan])


def test_convert_f64_to_f16_bitexact(handle, gpu):
    import torch
    rf = _rf()
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(5000) * 1e3, rng.standard_normal(5000) * 1e-5,
                        rng.standard_normal(5000)]).reshape(150, 100, order="F")
    a = to_dev(x, gpu)
    out = rf.colmajor_empty(150, 100, dtype=torch.float16, device=gpu)
    rf.convert(handle, a, out)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(out), orc.convert(x, orc.F64, orc.F16))


# ---------------------------------------------------------------- panel
def _panel_case(rng, prec, m, n):
    """diagonally dominant block column/row data in the storage precision"""
    l = rng.uniform(-1, 1, (m, n)).astype(np.float64)
    u = rng.uniform(-1, 1, (n, m)).astype(np.float64)
    for i in range(n):
        l[i, i] = 2.0 * n + rng.uniform(0, 1)
    d = l[:n, :n].copy()
    cast = {orc.F32: np.float32, orc.F64: np.float64}
    if prec == orc.F16:
        conv = lambda z: orc.convert(np.asfortranarray(z), orc.F64, orc.F16)
    else:
        conv = lambda z: np.asfortranarray(z.astype(cast[prec]))
    return conv(d), conv(l), conv(u)


@pytest.mark.parametrize("prec", ["f32", "f16", "f64"])
@pytest.mark.parametrize("ib,n,m", [(8, 8, 8), (8, 8, 300), (8, 5, 77), (16, 16, 1000),
                                    (32, 32, 513), (32, 19, 64), (8, 1, 9), (4, 4, 40)])
def test_panel_bitexact(handle, gpu, prec, ib, n, m):
    import torch
    rf = _rf()
    p = {"f32": orc.F32, "f16": orc.F16, "f64": orc.F64}[prec]
    rng = np.random.default_rng(ib * 1000 + n * 10 + m)
    d, l, u = _panel_case(rng, p, m, n)
    dl, ll, ul = to_dev(d, gpu, ld=ib), to_dev(l, gpu, ld=m + 2), to_dev(u, gpu, ld=n + 1)
    stat = torch.zeros(1, dtype=torch.int32, device=gpu)
    rf.lu_nopiv_panel(handle, ib, m, n, dl, ll, ul, stat)
    torch.cuda.synchronize()
    lref, uref, sref = orc.lu_nopiv_panel(p, ib, m, n, d, l, u)
    gl, gu = to_np(ll), to_np(ul)
    if p == orc.F32:
        gl, gu, lref, uref = (z.view(np.uint32) for z in (gl, gu, lref, uref))
    elif p == orc.F64:
        gl, gu, lref, uref = (z.view(np.uint64) for z in (gl, gu, lref, uref))
    assert np.array_equal(gl, lref)
    assert np.array_equal(gu[:, n:], uref[:, n:])
    assert int(stat.item()) == sref == n


# ---------------------------------------------------------------- gemm
GEMM_SHAPES = [(1, 1, 1), (37, 45, 19), (300, 257, 513), (1024, 8, 512), (8, 1000, 512),
               (520, 530, 64), (2048, 512, 1536), (64, 64, 4096)]


def _gemm_inputs(rng, m, n, k, half_in):
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    c = rng.uniform(-4, 4, (m, n)).astype(np.float32)
    if half_in:
        a, b = orc.round_to(orc.F16, a), orc.round_to(orc.F16, b)
    return np.asfortranarray(a), np.asfortranarray(b), np.asfortranarray(c)


@pytest.mark.parametrize("ctype,out", [("FP32", "f32"), ("TC32", "f32"), ("TC32", "f16"),
                                       ("BF16_TC32", "f32")])
@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_mfma_within_bound(handle, gpu, ctype, out, m, n, k):
    import torch
    rf = _rf()
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    half_in = ctype != "FP32"
    a, b, c = _gemm_inputs(rng, m, n, k, half_in)
    if out == "f16":
        c = orc.round_to(orc.F16, c)
    ct = getattr(rf, ctype)
    oc = {"FP32": orc.FP32, "TC32": orc.TC32, "BF16_TC32": orc.BF16TC32}[ctype]
    if half_in:
        da, db = to_dev(orc.f32_to_f16_bits(a), gpu, ld=m + 8), to_dev(orc.f32_to_f16_bits(b), gpu)
    else:
        da, db = to_dev(a, gpu, ld=m + 1), to_dev(b, gpu)
    dc = to_dev(orc.f32_to_f16_bits(c) if out == "f16" else c, gpu)
    rf.gemmex(handle, ct, da, db, dc, alpha=-1.0, beta=1.0)
    torch.cuda.synchronize()
    got = to_np(dc)
    got = orc.f16_bits_to_f32(got) if out == "f16" else got
    ref = orc.gemm_sub(oc, a, b, c, orc.F16 if out == "f16" else orc.F32)
    # operands as the compute type sees them
    opp = {"FP32": orc.F32, "TC32": orc.F16, "BF16_TC32": orc.BF16}[ctype]
    aa, bb = orc.round_to(opp, a).astype(np.float64), orc.round_to(opp, b).astype(np.float64)
    absab = np.abs(aa) @ np.abs(bb)
    u = 2.0**-24
    # both results are within gamma_{k+1} (|A||B| + |C|) of the exact value
    # (the GPU may accumulate the products onto C directly)
    gam = (k + 1) * u / (1 - (k + 1) * u)
    bound = 2 * gam * (absab + np.abs(c))
    if out == "f16":
        bound += 2 * 2.0**-11 * np.abs(ref)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound + 1e-30)


@pytest.mark.parametrize("tile", range(8))
@pytest.mark.parametrize("ctype", ["TC32", "BF16_TC32"])
@pytest.mark.parametrize("m,n,k", [(300, 257, 513), (512, 384, 256)])
def test_gemm_every_tile_within_bound(handle, gpu, tile, ctype, m, n, k):
    """each matrix-core tile shape (forced), aligned and ragged operands"""
    import torch
    rf = _rf()
    rng = np.random.default_rng(tile + m)
    a, b, c = _gemm_inputs(rng, m, n, k, True)
    oc = {"TC32": orc.TC32, "BF16_TC32": orc.BF16TC32}[ctype]
    da, db = to_dev(orc.f32_to_f16_bits(a), gpu), to_dev(orc.f32_to_f16_bits(b), gpu)
    dc = to_dev(c, gpu)
    handle.set_gemm_tile(tile)
    try:
        rf.gemmex(handle, getattr(rf, ctype), da, db, dc, alpha=-1.0, beta=1.0)
        torch.cuda.synchronize()
    finally:
        handle.set_gemm_tile(-1)
    got = to_np(dc)
    ref = orc.gemm_sub(oc, a, b, c, orc.F32)
    opp = orc.F16 if ctype == "TC32" else orc.BF16
    aa, bb = orc.round_to(opp, a).astype(np.float64), orc.round_to(opp, b).astype(np.float64)
    gam = (k + 1) * 2.0**-24 / (1 - (k + 1) * 2.0**-24)
    bound = 2 * gam * (np.abs(aa) @ np.abs(bb) + np.abs(c))
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound + 1e-30)


@pytest.mark.parametrize("m,n,k,out,split,storage", [
    (512, 512, 256, "f32", 0, "f16"),
    (768, 512, 1536, "f32", 0, "f16"),    # 48 slabs: the steady-state loop
    (256, 1024, 64, "f32", 0, "f16"),     # fewer slabs than the prefetch depth
    (256, 256, 32, "f32", 0, "f16"),      # a single slab
    (512, 256, 128, "f16", 0, "f16"),     # fp16 output
    (1024, 512, 2048, "f32", 4, "f16"),   # split-K partials
    (512, 256, 1024, "f16", 3, "f16"),    # ... three slices, fp16 output
    (512, 768, 512, "f32", 0, "bf16"),    # bf16 storage
])
def test_gemm_large_tile_within_bound(handle, gpu, m, n, k, out, split, storage):
    """the 256 x 256 tile kernel (LDS-DMA operand slabs, swizzled layouts)
    on shapes it accepts, with padded leading dimensions"""
    import torch
    rf = _rf()
    rng = np.random.default_rng(m + 3 * n + k)
    a, b, c = _gemm_inputs(rng, m, n, k, True)
    if storage == "bf16":
        a, b = orc.round_to(orc.BF16, a), orc.round_to(orc.BF16, b)
        bits = lambda x: (np.asfortranarray(x).view(np.uint32) >> 16).astype(np.uint16)
        da = to_dev(bits(a), gpu, ld=m + 8).view(torch.bfloat16)
        db = to_dev(bits(b), gpu, ld=k + 16).view(torch.bfloat16)
        ct, oc, opp = rf.BF16_TC32, orc.BF16TC32, orc.BF16
    else:
        da = to_dev(orc.f32_to_f16_bits(a), gpu, ld=m + 8)
        db = to_dev(orc.f32_to_f16_bits(b), gpu, ld=k + 16)
        ct, oc, opp = rf.TC32, orc.TC32, orc.F16
    if out == "f16":
        c = orc.round_to(orc.F16, c)
    c_dev = orc.f32_to_f16_bits(c) if out == "f16" else c
    runs = []
    handle.set_gemm_tile(4)
    handle.set_tuning(0, split)
    try:
        # the split-K partials are added in slice order: a second run gives
        # the same bits
        for _ in range(2 if split > 1 else 1):
            dc = to_dev(c_dev, gpu, ld=m + 4)
            rf.gemmex(handle, ct, da, db, dc, alpha=-1.0, beta=1.0)
            torch.cuda.synchronize()
            runs.append(to_np(dc))
    finally:
        handle.set_gemm_tile(-1)
        handle.set_tuning(0, 0)
    if len(runs) == 2:
        assert runs[0].tobytes() == runs[1].tobytes()
    got = runs[0]
    got = orc.f16_bits_to_f32(got) if out == "f16" else got
    ref = orc.gemm_sub(oc, a, b, c, orc.F16 if out == "f16" else orc.F32)
    aa, bb = orc.round_to(opp, a).astype(np.float64), orc.round_to(opp, b).astype(np.float64)
    gam = (k + 1) * 2.0**-24 / (1 - (k + 1) * 2.0**-24)
    bound = 2 * gam * (np.abs(aa) @ np.abs(bb) + np.abs(c))
    if out == "f16":
        bound += 2 * 2.0**-11 * np.abs(ref)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound + 1e-30)


@pytest.mark.parametrize("ctype", ["FP16", "TC16"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 45, 19), (130, 70, 200), (64, 64, 257)])
def test_gemm_h16_bitexact(handle, gpu, ctype, m, n, k):
    import torch
    rf = _rf()
    rng = np.random.default_rng(3 * m + n + k)
    a, b, c = _gemm_inputs(rng, m, n, k, True)
    a, b = a * 0.25, b * 0.25  # keep fp16 partial sums well inside range
    c = orc.round_to(orc.F16, c)
    da, db = to_dev(orc.f32_to_f16_bits(a), gpu), to_dev(orc.f32_to_f16_bits(b), gpu)
    dc = to_dev(orc.f32_to_f16_bits(c), gpu)
    rf.gemmex(handle, getattr(rf, ctype), da, db, dc)
    torch.cuda.synchronize()
    got = to_np(dc)
    ref = orc.f32_to_f16_bits(orc.gemm_sub(orc.FP16, a, b, c, orc.F16))
    assert np.array_equal(got, ref)


def test_gemm_split_k_matches_unsplit(handle, gpu):
    """split-K partial sums are added in slice order: forcing 4 slices gives
    the same result as the default heuristic within the gamma_k bound, and
    repeated runs are bitwise reproducible."""
    import torch
    rf = _rf()
    rng = np.random.default_rng(5)
    m, n, k = 256, 64, 8192
    a, b, c = _gemm_inputs(rng, m, n, k, True)
    da, db = to_dev(orc.f32_to_f16_bits(a), gpu), to_dev(orc.f32_to_f16_bits(b), gpu)
    outs = []
    for split in (1, 4, 4):
        handle.set_tuning(0, split)
        dc = to_dev(c, gpu)
        rf.gemmex(handle, rf.TC32, da, db, dc)
        torch.cuda.synchronize()
        outs.append(to_np(dc))
    handle.set_tuning(0, 0)
    assert np.array_equal(outs[1].view(np.uint32), outs[2].view(np.uint32))
    absab = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64))
    gam = k * 2.0**-24 / (1 - k * 2.0**-24)
    assert np.all(np.abs(outs[0].astype(np.float64) - outs[1]) <= 4 * gam * absab + 1e-6)
