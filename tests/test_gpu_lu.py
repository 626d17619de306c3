"""GPU LU drivers (C ABI) against the oracle's restatement of the reference.

For a precision tuple the GPU and the oracle perform the same operations
with the same rounding points; they differ only in the summation order inside
the GEMM updates (matrix cores vs. k-ordered fma). Tolerance (written in the
test): the distance between GPU and oracle factors must not exceed
8 * max(distance(oracle, FP64 LU), unit roundoff of the storage precision),
i.e. the GPU result is as close to the oracle as the oracle is to the exact
factors. Each factor is normalised by its own magnitude.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import factor_distance, to_dev, to_np

pytestmark = pytest.mark.gpu

U = {orc.F32: 2.0**-24, orc.F16: 2.0**-11}


def _rf():
    import remifa_amd as rf
    return rf


def _oc(rf_code):
    return {0: orc.F32, 1: orc.F16}[rf_code]


def _ocomp(c):
    return c  # the compute-type codes are shared by the C ABI and the oracle


def _run_gpu_helper(handle, gpu, name, a32, nb):
    import torch
    rf = _rf()
    row = [r for r in rf.REFERENCE_HELPERS if r[0] == name][0]
    _, with_cpy, W = row[:3]
    n = a32.shape[0]
    ld = n + 5
    if W == rf.R_32F:
        da = to_dev(a32, gpu, ld=ld)
    else:
        da = to_dev(orc.f32_to_f16_bits(a32), gpu, ld=ld)
    info = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    fn = getattr(rf, name)
    if with_cpy:
        cpy = rf.colmajor_empty(n, n, dtype=torch.float16, device=gpu, ld=n + 3)
        fn(handle, n, n, nb, da, ld, info, cpy)
    else:
        fn(handle, n, n, nb, da, ld, info)
    torch.cuda.synchronize()
    out = to_np(da)
    return (orc.f16_bits_to_f32(out) if W == rf.R_16F else out), int(info.item())


def _oracle(row, a32, nb, algo="ll"):
    _, _, W, I, O, OP, IP, OF, IF = row
    c = orc.LuCfg(_oc(W), I, O, _oc(OP), _oc(IP), _oc(OF), _oc(IF), 8)
    a = a32 if c.W == orc.F32 else orc.f32_to_f16_bits(a32)
    f = orc.lu_nopiv(algo, c, a, nb)
    return (orc.f16_bits_to_f32(f) if c.W == orc.F16 else f), c


def _check(gpu_f, ora_f, f64, W):
    e_o = factor_distance(ora_f, f64)
    d = factor_distance(gpu_f, ora_f)
    tol = 8 * max(e_o, U[W])
    assert d <= tol, f"GPU vs oracle {d:.3e} > tol {tol:.3e} (oracle vs FP64 {e_o:.3e})"
    e_g = factor_distance(gpu_f, f64)
    assert e_g <= 8 * max(e_o, U[W])


HELPERS = ["lu_nopiv_ll_f32", "lu_nopiv_ll_f32_fp32xtc32", "lu_nopiv_ll_f32_tc32xtc32",
           "lu_nopiv_ll_f16", "lu_nopiv_ll_f16_tc16xtc16", "lu_nopiv_ll_f16_tc32xtc32",
           "lu_nopiv_ll_f16_f16xtc32", "lu_nopiv_ll_f16_f16xtc32_fp32", "lu_nopiv_ll_f16_f32xt32",
           "lu_nopiv_ll_f16_f16xtc16", "lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32",
           "lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32"]


@pytest.mark.parametrize("name", HELPERS)
@pytest.mark.parametrize("n,nb,alpha", [(256, 64, 1.5), (200, 48, 1.1)])
def test_reference_helper_matches_oracle(handle, gpu, name, n, nb, alpha):
    rf = _rf()
    row = [r for r in rf.REFERENCE_HELPERS if r[0] == name][0]
    a32 = orc.gen_diagdom_alpha(n, 11, alpha)
    if row[2] == rf.R_16F:
        a_used = orc.round_to(orc.F16, a32)
    else:
        a_used = a32
    f64 = orc.lu_f64(a_used)
    g, info = _run_gpu_helper(handle, gpu, name, a32, nb)
    o, c = _oracle(row, a32, nb)
    assert info == 0
    _check(g, o, f64, c.W)


@pytest.mark.parametrize("W,I,O,nb", [(0, 0, 0, 64), (0, 0, 2, 64), (0, 2, 2, 32), (1, 2, 2, 64),
                                      (1, 1, 1, 64), (1, 3, 3, 32), (1, 1, 2, 64), (1, 4, 4, 64),
                                      (1, 2, 4, 64)])
def test_rl_matches_oracle(handle, gpu, W, I, O, nb):
    import torch
    rf = _rf()
    n = 256
    a32 = orc.gen_diagdom_alpha(n, 5, 1.5)
    ow = _oc(W)
    c = orc.LuCfg(ow, I, O, ow, ow, ow, ow, 8)
    a_host = a32 if ow == orc.F32 else orc.f32_to_f16_bits(a32)
    da = to_dev(a_host, gpu, ld=n + 1)
    cfg = rf.api.make_config(W=W, I=I, O=O, ib=8)
    info = rf.lu_nopiv(handle, da, nb, algo=rf.RL, cfg=cfg)
    torch.cuda.synchronize()
    g = to_np(da)
    g = orc.f16_bits_to_f32(g) if ow == orc.F16 else g
    o = orc.lu_nopiv("rl", c, a_host, nb)
    o = orc.f16_bits_to_f32(o) if ow == orc.F16 else o
    f64 = orc.lu_f64(a32 if ow == orc.F32 else orc.round_to(orc.F16, a32))
    assert info == 0
    _check(g, o, f64, ow)


def test_c1_plain_fp32_ll_2048(handle, gpu):
    """Config 1: plain FP32 LL, nb=256, ib=8, n=2048, a_ii = 1.5 sum|a_ij|:
    factor error ||A - LU|| / ||A|| and solve error against FP64."""
    import torch
    rf = _rf()
    n, nb = 2048, 256
    a32 = orc.gen_diagdom_alpha(n, 1, 1.5)
    g, info = _run_gpu_helper(handle, gpu, "lu_nopiv_ll_f32", a32, nb)
    assert info == 0
    L, Uu = np.tril(g.astype(np.float64), -1) + np.eye(n), np.triu(g.astype(np.float64))
    rel = np.linalg.norm(a32.astype(np.float64) - L @ Uu) / np.linalg.norm(a32.astype(np.float64))
    assert rel < 1e-6
    # check solve of the reference harness (x = 1, b = A x, fp32 getrs with
    # identity pivots): the GPU factors solve as well as the oracle's
    b = a32.astype(np.float64) @ np.ones(n)
    o, _ = _oracle([r for r in rf.REFERENCE_HELPERS if r[0] == "lu_nopiv_ll_f32"][0], a32, nb)
    x = orc.getrs_f32(g, b.astype(np.float32))[:, 0].astype(np.float64)
    xo = orc.getrs_f32(o, b.astype(np.float32))[:, 0].astype(np.float64)
    e_g, e_o = orc.backward_error(a32, b, x), orc.backward_error(a32, b, xo)
    assert e_g <= 4 * max(e_o, 2.0**-24)
    _check(g, o, orc.lu_f64(a32), orc.F32)


def test_info_reports_zero_pivot(handle, gpu):
    import torch
    rf = _rf()
    n = 64
    a = np.asfortranarray(np.eye(n, dtype=np.float32))
    a[10, 10] = 0.0
    da = to_dev(a, gpu)
    info = rf.lu_nopiv(handle, da, 16, algo=rf.LL, cfg=rf.api.make_config())
    assert info == 11


def test_bad_arguments_raise(handle, gpu):
    import torch
    rf = _rf()
    a = rf.colmajor_empty(8, 8, device=gpu)
    with pytest.raises(rf.RemifaError) as e:
        rf.lu_nopiv(handle, a, 0, algo=rf.LL, cfg=rf.api.make_config())
    assert e.value.status == -6
    info = torch.zeros(1, dtype=torch.int32, device=gpu)
    with pytest.raises(rf.RemifaError) as e:
        rf.lu_nopiv_ll_f32_fp32xtc32(handle, 8, 8, 4, a, 8, info, None)
    assert e.value.status == -8


@pytest.mark.parametrize("OP,n,nb,ib", [("f32", 320, 64, 8), ("f32", 500, 96, 32), ("f16", 256, 64, 8)])
def test_ll_bf16_outer_updates_match_oracle(handle, gpu, OP, n, nb, ib):
    """LL with BF16_TC32 outer updates: fp32 inner buffers (outer operands
    from the LU's bf16 mirror of its factors) and in place (operands rounded
    while staged), against the oracle's restatement."""
    import torch
    rf = _rf()
    a32 = orc.gen_diagdom_alpha(n, 23, 1.2)
    a16 = orc.f32_to_f16_bits(a32)
    op_rf, op_or = (rf.R_32F, orc.F32) if OP == "f32" else (rf.R_16F, orc.F16)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.BF16_TC32, OP=op_rf, ib=ib)
    da = to_dev(a16, gpu, ld=n + 3)
    assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
    torch.cuda.synchronize()
    g = orc.f16_bits_to_f32(to_np(da))
    o = orc.f16_bits_to_f32(orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.BF16TC32,
                                                        OP=op_or, ib=ib), a16, nb))
    _check(g, o, orc.lu_f64(orc.round_to(orc.F16, a32)), orc.F16)


@pytest.mark.parametrize("O", ["tc32", "bf16"])
def test_ll_lookahead_matches_oracle_and_sequential(handle, gpu, O):
    """LL with fp32 inner buffers: the lookahead (block k+1's update by the
    blocks before k on a second stream, block k's part afterwards) computes
    the same products as one update per block, summed in two parts. Both
    schedules against the oracle, with a ragged last block, in a
    caller-owned workspace of exactly remifa_lu_workspace_size bytes."""
    import torch
    rf = _rf()
    n, nb, ib = 700, 128, 32
    a32 = orc.gen_diagdom_alpha(n, 31, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    oc_o, rf_o = (orc.TC32, rf.TC32) if O == "tc32" else (orc.BF16TC32, rf.BF16_TC32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf_o, OP=rf.R_32F, ib=ib)
    nbytes = rf.lu_workspace_size(handle, rf.LL, cfg, n, n, nb)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu)
    handle.set_workspace(ws)
    out = {}
    try:
        # (on / off, CUs masked off the second stream)
        for on, rsv in ((1, 32), (1, 0), (0, 32)):
            handle.set_lookahead(on, rsv)
            da = to_dev(a16, gpu, ld=n + 3)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            out[(on, rsv)] = orc.f16_bits_to_f32(to_np(da))
    finally:
        handle.set_lookahead(1, 0)
        handle.set_workspace(None)
    o = orc.f16_bits_to_f32(orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=oc_o,
                                                        OP=orc.F32, ib=ib), a16, nb))
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    for f in out.values():
        _check(f, o, f64, orc.F16)
    # the masked and unmasked runs differ only in where the workgroups ran
    assert np.array_equal(out[(1, 32)], out[(1, 0)])
    assert factor_distance(out[(1, 0)], out[(0, 32)]) <= 8 * U[orc.F16]


def test_set_lookahead_rejects_bad_arguments(handle, gpu):
    """remifa_set_lookahead: on must be 0 / 1; reserve_cus a multiple of 8,
    at most half the CUs (status -2 / -3, the knob unchanged)"""
    rf = _rf()
    from remifa_amd._lib import lib
    h = handle._h
    assert lib().remifa_set_lookahead(h, 2, 0) == -2
    assert lib().remifa_set_lookahead(h, 1, 12) == -3
    assert lib().remifa_set_lookahead(h, 1, -8) == -3
    assert lib().remifa_set_lookahead(h, 1, 4096) == -3
    assert lib().remifa_set_lookahead(None, 1, 0) == -1
    assert lib().remifa_set_lookahead(h, 1, 0) == 0
    with pytest.raises(rf.RemifaError):
        handle.set_lookahead(True, 7)
