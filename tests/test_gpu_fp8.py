"""FP8_TC32 outer updates (an MI355X addition with no reference code; the
oracle's FP8TC32 restatement is pinned in test_oracle.py by hand-derived
codes and by torch's float8_e4m3fn cast).

* the quantizer (remifa_quantize_fp8) against the oracle's operand format,
  codes and scale exponents bit for bit;
* the scaled matrix-core GEMM against the oracle's FP8TC32 update: both sum
  the same exactly representable products of dequantized codes, so they
  differ only by fp32 summation order, bounded by 2 k 2^-24 sum|a_q||b_q|;
* LL (fp32 inner buffers and in place) and RL LU with FP8 outer updates
  against the oracle, with the tolerance of the other LU tests;
* FP64-residual refinement from FP8-updated factors to a backward error of
  1e-14.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import factor_distance, to_dev, to_np

pytestmark = pytest.mark.gpu


def _operand(rng, m, k, spread=10):
    """rows of very different magnitudes and some all-zero 32-blocks"""
    a = rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-spread, spread, (m, 1)))
    a = a.astype(np.float32)
    if k > 64:
        a[:, 32:64] = 0.0
    a[min(3, m - 1), :] = 0.0
    return np.asfortranarray(a)


def _dequant(codes, sexp):
    import torch
    v = torch.from_numpy(np.ascontiguousarray(codes)).view(torch.float8_e4m3fn).double().numpy()
    return v * np.exp2(sexp.repeat(32, axis=1)[:, : codes.shape[1]].astype(np.float64))


@pytest.mark.parametrize("rows", [True, False])
@pytest.mark.parametrize("src", ["f16", "f32"])
def test_quantize_matches_oracle(handle, gpu, rows, src):
    import remifa_amd as rf
    rng = np.random.default_rng(3)
    m, k = 77, 200  # k: six whole 32-blocks and a partial one
    a = _operand(rng, m, k, spread=14)
    if src == "f16":
        a = orc.round_to(orc.F16, a)
    x = a if rows else np.asfortranarray(a.T)
    host = orc.f32_to_f16_bits(x) if src == "f16" else x
    q, s = rf.quantize_fp8(handle, to_dev(host, gpu, ld=x.shape[0] + 3), rows=rows)
    codes, sexp = orc.quantise_e4m3(x, rows=rows)
    qn, sn = q.cpu().numpy(), s.cpu().numpy()
    nkb = (k + 31) // 32
    assert np.array_equal(qn[:, :k], codes)
    assert not qn[:, k:].any()  # the partial block is padded with zero codes
    assert np.array_equal(sn[:, :nkb].astype(np.int64) - 127, sexp)


@pytest.mark.parametrize("m,n,k,split", [(256, 256, 256, 0), (300, 200, 96, 0),
                                         (130, 515, 1056, 0), (512, 384, 2048, 3),
                                         (64, 1000, 4096, 0)])
def test_gemm_fp8_matches_oracle(handle, gpu, m, n, k, split):
    import torch
    import remifa_amd as rf
    rng = np.random.default_rng(m + n + k)
    a = orc.round_to(orc.F16, _operand(rng, m, k))
    b = orc.round_to(orc.F16, np.asfortranarray(_operand(rng, n, k).T))
    c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    dc = to_dev(c, gpu, ld=m + 2)
    handle.set_tuning(0, split)
    try:
        rf.gemmex(handle, rf.FP8_TC32, to_dev(orc.f32_to_f16_bits(a), gpu, ld=m + 1),
                  to_dev(orc.f32_to_f16_bits(b), gpu, ld=k + 16), dc)
        torch.cuda.synchronize()
    finally:
        handle.set_tuning(0, 0)
    got = to_np(dc).astype(np.float64)
    want = orc.gemm_sub(orc.FP8TC32, a, b, c).astype(np.float64)
    qa = _dequant(*orc.quantise_e4m3(a, rows=True))
    qb = _dequant(*orc.quantise_e4m3(b, rows=False))
    absprod = np.abs(qa) @ np.abs(qb).T
    tol = 2 * k * 2.0**-24 * absprod + 2.0**-23 * np.abs(c)
    assert np.all(np.abs(got - want) <= tol), np.max(np.abs(got - want) / (tol + 1e-300))
    # the products themselves: against the exactly summed dequantized operands
    exact = c - qa @ qb.T
    assert np.all(np.abs(got - exact) <= tol)


def test_gemm_fp8_from_quantized_operands_f16_output(handle, gpu):
    """remifa_quantize_fp8 + remifa_gemm_fp8 with an fp16 C (alpha, beta general)"""
    import torch
    import remifa_amd as rf
    rng = np.random.default_rng(8)
    m, n, k = 200, 136, 320
    a = _operand(rng, m, k, spread=2)  # results within the fp16 range
    b = np.asfortranarray(_operand(rng, n, k, spread=2).T)
    c = orc.round_to(orc.F16, np.asfortranarray(rng.uniform(-4, 4, (m, n)).astype(np.float32)))
    qa, sa = rf.quantize_fp8(handle, to_dev(a, gpu), rows=True)
    qb, sb = rf.quantize_fp8(handle, to_dev(b, gpu), rows=False)
    dc = to_dev(orc.f32_to_f16_bits(c), gpu)
    rf.gemm_fp8(handle, qa, sa, qb, sb, k, dc, alpha=0.5, beta=-2.0)
    torch.cuda.synchronize()
    got = orc.f16_bits_to_f32(to_np(dc)).astype(np.float64)
    pa = _dequant(*orc.quantise_e4m3(a, rows=True))
    pb = _dequant(*orc.quantise_e4m3(b, rows=False))
    want = 0.5 * (pa @ pb.T) - 2.0 * c
    tol = 2.0**-11 * np.abs(want) + 2 * k * 2.0**-24 * (np.abs(pa) @ np.abs(pb).T) + 2.0**-24
    assert np.all(np.abs(got - want) <= tol)


@pytest.mark.parametrize("split", [0, 4])
@pytest.mark.parametrize("algo,OP,n,nb,ib", [("ll", "f32", 320, 64, 8), ("ll", "f16", 256, 64, 8),
                                             ("ll", "f32", 500, 96, 32), ("rl", "f16", 256, 64, 8),
                                             ("ll", "f32", 600, 128, 32)])
def test_lu_fp8_outer_matches_oracle(handle, gpu, algo, OP, n, nb, ib, split):
    """(LL with fp32 buffers and nb = 128: the lookahead schedule, block k's
    part of each update taken from K offset (k - 1) nb of the e4m3 mirrors;
    split = 4 forces split-K, so that the C-input updates go through the
    split-K reduction). Entrywise parity with the oracle (tests/helpers.py),
    then the error level against the FP64 LU."""
    import torch
    import remifa_amd as rf
    from helpers import assert_parity
    a32 = orc.gen_diagdom_alpha(n, 17, 1.2)
    a16 = orc.f32_to_f16_bits(a32)
    op_rf = rf.R_32F if OP == "f32" else rf.R_16F
    op_or = orc.F32 if OP == "f32" else orc.F16
    if algo == "rl":
        cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.FP8_TC32, ib=ib)
        oc = orc.cfg(W=orc.F16, I=orc.TC32, O=orc.FP8TC32, ib=ib)
    else:
        cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.FP8_TC32, OP=op_rf, ib=ib)
        oc = orc.cfg(W=orc.F16, I=orc.TC32, O=orc.FP8TC32, OP=op_or, ib=ib)
    da = to_dev(a16, gpu, ld=n + 3)
    try:
        handle.set_tuning(0, split)
        info = rf.lu_nopiv(handle, da, nb, algo=rf.LL if algo == "ll" else rf.RL, cfg=cfg)
        torch.cuda.synchronize()
    finally:
        handle.set_tuning(0, 0)
    assert info == 0
    graw = to_np(da)
    oraw = orc.lu_nopiv(algo, oc, a16, nb)
    assert_parity(graw, oraw, f"fp8 {algo} OP={OP} n={n} nb={nb} split={split}")
    g, o = orc.f16_bits_to_f32(graw), orc.f16_bits_to_f32(oraw)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    e_o = factor_distance(o, f64)
    e_g = factor_distance(g, f64)
    assert e_g <= 8 * max(e_o, 2.0**-11), f"GPU vs FP64 {e_g:.3e} (oracle vs FP64 {e_o:.3e})"


def test_fp8_update_ir_reaches_1e14(handle, gpu):
    import torch
    import remifa_amd as rf
    n, nb = 1536, 128
    dev = torch.device(gpu)
    a = rf.colmajor_empty(n, n, device=dev)
    rf.gen_diagdom(handle, a, seed=5, alpha=1.05)
    a16 = rf.colmajor_empty(n, n, dtype=torch.float16, device=dev)
    rf.convert(handle, a, a16)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.FP8_TC32, OP=rf.R_32F, ib=32)
    assert rf.lu_nopiv(handle, a16, nb, algo=rf.LL, cfg=cfg) == 0
    bf = rf.colmajor_empty(n, 1, device=dev)
    rf.gen_uniform(handle, bf, seed=6)
    b = torch.zeros((1, n), dtype=torch.float64, device=dev).t()
    b.copy_(bf)
    x = torch.zeros((1, n), dtype=torch.float64, device=dev).t()
    it, be, hist = rf.ir_solve(handle, a, a16, b, x, tol=1e-14, maxit=30)
    assert it >= 0 and be <= 1e-14, hist
    host = orc.backward_error(to_np(a), b.cpu().numpy(), x.cpu().numpy())
    assert host <= 1e-14
