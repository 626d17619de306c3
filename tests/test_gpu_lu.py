"""GPU LU drivers (C ABI) against the oracle's restatement of the reference.

For a precision tuple the GPU and the oracle perform the same operations
with the same rounding points; they differ only in the summation order inside
the GEMM updates (matrix cores vs. k-ordered fma). Two assertions per case
(tolerances written in tests/helpers.py):
 1. entrywise, in the storage format: at least F16_MIN_EQUAL (fp16) /
    F32_MIN_EQUAL (fp32) of the entries bit-identical to the oracle's and
    none more than F16_MAX_ULP / F32_MAX_ULP units in the last place away
    (a misplaced rounding point moves entries by tens to thousands of ulps:
    tests/test_oracle.py::test_parity_rule_rejects_planted_rounding_mistakes);
 2. error level: the GPU factors are as close to the FP64 LU as the
    oracle's: distance(GPU, FP64) <= 8 max(distance(oracle, FP64), u_W),
    each factor normalised by its own magnitude.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity, factor_distance, to_dev, to_np

pytestmark = pytest.mark.gpu

U = {orc.F32: 2.0**-24, orc.F16: 2.0**-11}


def _rf():
    import remifa_amd as rf
    return rf


def _oc(rf_code):
    return {0: orc.F32, 1: orc.F16}[rf_code]


def _f(raw):
    """storage-format factors (fp16 bits or fp32) as float32 values"""
    return orc.f16_bits_to_f32(raw) if raw.dtype == np.uint16 else raw


def _run_gpu_helper(handle, gpu, name, a32, nb):
    """one of the 12 reference helpers on the GPU; returns the factors in
    storage format (fp16 bits or fp32) and info"""
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
    return to_np(da), int(info.item())


def _oracle(row, a32, nb, algo="ll"):
    _, _, W, I, O, OP, IP, OF, IF = row
    c = orc.LuCfg(_oc(W), I, O, _oc(OP), _oc(IP), _oc(OF), _oc(IF), 8)
    a = a32 if c.W == orc.F32 else orc.f32_to_f16_bits(a32)
    return orc.lu_nopiv(algo, c, a, nb), c


def _check(gpu_raw, ora_raw, f64, W, label=""):
    """entrywise parity in storage format, then the error level"""
    assert_parity(gpu_raw, ora_raw, label)
    g, o = _f(gpu_raw), _f(ora_raw)
    e_o = factor_distance(o, f64)
    e_g = factor_distance(g, f64)
    assert e_g <= 8 * max(e_o, U[W]), f"{label}: GPU vs FP64 {e_g:.3e}, oracle vs FP64 {e_o:.3e}"


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
    _check(g, o, f64, c.W, f"{name} n={n} nb={nb} alpha={alpha}")


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
    o = orc.lu_nopiv("rl", c, a_host, nb)
    f64 = orc.lu_f64(a32 if ow == orc.F32 else orc.round_to(orc.F16, a32))
    assert info == 0
    _check(g, o, f64, ow, f"rl W={W} I={I} O={O} nb={nb}")


def test_c1_plain_fp32_ll_2048(handle, gpu):
    """Config 1: plain FP32 LL, nb=256, ib=8, n=2048, a_ii = 1.5 sum|a_ij|:
    factor error ||A - LU|| / ||A|| and solve error against FP64."""
    import torch
    rf = _rf()
    n, nb = 2048, 256
    a32 = orc.gen_diagdom_alpha(n, 1, 1.5)
    g, info = _run_gpu_helper(handle, gpu, "lu_nopiv_ll_f32", a32, nb)
    assert info == 0
    # factor error on the GPU (tests/errors.hxx:251-346) against the oracle's
    fe = rf.factor_error(handle, to_dev(a32, gpu), to_dev(g, gpu))
    fo = orc.factor_error(a32, g)
    assert fe["factor_cwerr"] == pytest.approx(fo, rel=1e-9)
    assert fe["factor_cwerr"] < 64 * 2.0**-24
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
    assert rel == pytest.approx(fe["factor_rel_fro"], rel=1e-6)
    _check(g, o, orc.lu_f64(a32), orc.F32, "C1 lu_nopiv_ll_f32 n=2048 nb=256")


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
    g = to_np(da)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.BF16TC32, OP=op_or, ib=ib),
                     a16, nb)
    _check(g, o, orc.lu_f64(orc.round_to(orc.F16, a32)), orc.F16, f"ll bf16 OP={OP} n={n}")


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
            out[(on, rsv)] = to_np(da)
    finally:
        handle.set_lookahead(1, 0)
        handle.set_workspace(None)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=oc_o, OP=orc.F32, ib=ib), a16, nb)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    for k, f in out.items():
        _check(f, o, f64, orc.F16, f"lookahead {O} (on, reserve)={k}")
    # the masked and unmasked runs differ only in where the workgroups ran
    assert np.array_equal(out[(1, 32)], out[(1, 0)])
    assert_parity(out[(1, 0)], out[(0, 32)], "lookahead on vs off")


def test_set_lookahead_rejects_bad_arguments(handle, gpu):
    """remifa_set_lookahead: on must be 0 / 1; reserve_cus a multiple of 8 in
    [0, 64] (status -2 / -3); a rejected call leaves both settings as they
    were (read back with remifa_get_lookahead), and a lookahead LU still
    runs afterwards"""
    import ctypes as C
    import torch
    rf = _rf()
    from remifa_amd._lib import lib
    h = handle._h
    on, rsv = C.c_int(), C.c_int()
    assert lib().remifa_set_lookahead(h, 1, 16) == 0
    for bad_on, bad_rsv, st in ((2, 0, -2), (1, 12, -3), (1, -8, -3), (1, 72, -3), (1, 128, -3),
                                (0, 4096, -3)):
        assert lib().remifa_set_lookahead(h, bad_on, bad_rsv) == st
        assert lib().remifa_get_lookahead(h, C.byref(on), C.byref(rsv)) == 0
        assert (on.value, rsv.value) == (1, 16)
    assert lib().remifa_set_lookahead(None, 1, 0) == -1
    assert lib().remifa_get_lookahead(None, None, None) == -1
    assert lib().remifa_set_lookahead(h, 1, 0) == 0
    assert lib().remifa_get_lookahead(h, C.byref(on), C.byref(rsv)) == 0
    assert (on.value, rsv.value) == (1, 0)
    with pytest.raises(rf.RemifaError):
        handle.set_lookahead(True, 7)
    # the second stream is still used: the lookahead layout factors as before
    n, nb = 384, 128
    a16 = orc.f32_to_f16_bits(orc.gen_diagdom_alpha(n, 3, 1.2))
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    da = to_dev(a16, gpu)
    assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("O", ["tc32", "bf16", "fp8"])
def test_headline_schedule_matches_oracle(handle, gpu, O):
    """The headline configuration's schedule at n = 4096: fp16 storage, fp32
    inner buffers and panels (src/lu_nopiv_ll.cxx:618), nb = 512, ib = 32,
    lookahead on (eight outer blocks: seven second-stream updates, split-K
    chosen by the cost model), TC32 / BF16 / FP8 outer updates; and the same
    with split-K forced to 4 slices (the C-input reduction path). Both runs
    against the oracle, and against each other."""
    import torch
    rf = _rf()
    n, nb, ib = 4096, 512, 32
    a32 = orc.gen_diagdom_alpha(n, 2024, 1.05)
    a16 = orc.f32_to_f16_bits(a32)
    oc_o, rf_o = {"tc32": (orc.TC32, rf.TC32), "bf16": (orc.BF16TC32, rf.BF16_TC32),
                  "fp8": (orc.FP8TC32, rf.FP8_TC32)}[O]
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf_o, OP=rf.R_32F, ib=ib)
    out = {}
    try:
        handle.set_lookahead(True, 0)
        for split in (0, 4):
            handle.set_tuning(0, split)
            da = to_dev(a16, gpu)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            out[split] = to_np(da)
    finally:
        handle.set_tuning(0, 0)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=oc_o, OP=orc.F32, ib=ib), a16, nb)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    for split, f in out.items():
        _check(f, o, f64, orc.F16, f"headline n={n} nb={nb} ib={ib} O={O} split={split}")
    assert_parity(out[4], out[0], f"headline O={O} split 4 vs auto")


@pytest.mark.parametrize("n,nb,ld_pad", [(1000, 256, 4), (1536, 512, 0), (640, 192, 8),
                                         (777, 128, 3)])
def test_block_kernel_matches_slices_and_oracle(handle, gpu, n, nb, ld_pad):
    """The one-launch inner factorization of an outer block (blockll.hip:
    ib = 32, TC32 inner updates on fp32 buffers, outer blocks of a multiple
    of 64 columns, fp16 columns 8-byte aligned) against the oracle, and
    against the slice kernels it replaces (panel_width = 64 selects them).
    Cases: a ragged last block (1000 = 3 x 256 + 232 and 777 = 6 x 128 + 9
    take the slice kernels there), units cut by the matrix edge, nb = 192
    (three slices), and a leading dimension that is not a multiple of 4
    (every block on the slice kernels)."""
    import torch
    rf = _rf()
    a32 = orc.gen_diagdom_alpha(n, 7 + n, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    out = {}
    try:
        for pw in (0, 64):
            handle.set_tuning(pw, 0)
            da = to_dev(a16, gpu, ld=n + ld_pad)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            out[pw] = to_np(da)
    finally:
        handle.set_tuning(0, 0)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=32), a16, nb)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    _check(out[0], o, f64, orc.F16, f"block kernel n={n} nb={nb}")
    _check(out[64], o, f64, orc.F16, f"slice kernels n={n} nb={nb}")
    assert_parity(out[0], out[64], f"block kernel vs slice kernels n={n} nb={nb}")


@pytest.mark.parametrize("n,nb,ld_pad,upw,lookahead,oracle",
                         [(1536, 512, 0, 0, True, True), (1000, 256, 4, 1, True, True),
                          (640, 192, 8, 3, False, True), (777, 128, 0, 64, True, True),
                          (4096, 512, 0, 0, True, False), (3000, 512, 0, 5, True, False),
                          (640, 64, 0, 0, True, True), (512, 512, 0, 0, False, True)])
def test_block_kernels_one_launch_bit_identical(handle, gpu, n, nb, ld_pad, upw, lookahead,
                                                oracle):
    """The block kernels' one-launch form (the diagonal role publishing each
    slice, the far units following it in the same grid) against the
    two-launch form: bit-identical factors, with the helper count varied
    (far units per wavefront 1 .. 64; 0 = automatic), ragged blocks and
    units cut by the matrix edge, one slice per block (nb = 64: no
    diagonal-block units for the helpers) and a single outer block (no far
    units); against the oracle where it finishes in seconds."""
    import torch
    rf = _rf()
    a32 = orc.gen_diagdom_alpha(n, 13 + n, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    out = {}
    try:
        handle.set_lookahead(lookahead, 0)
        for one in (True, False, True):
            handle.set_block_kernels(one, upw)
            da = to_dev(a16, gpu, ld=n + ld_pad)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            f = to_np(da)
            if one in out:  # a second one-launch run: the same bits again
                assert f.tobytes() == out[one].tobytes(), f"one-launch rerun differs n={n}"
            out[one] = f
    finally:
        handle.set_block_kernels(True, 0)
        handle.set_lookahead(True, 0)
    assert out[True].tobytes() == out[False].tobytes(), f"one vs two launches differ n={n} nb={nb}"
    if oracle:
        o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=32),
                         a16, nb)
        f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
        _check(out[True], o, f64, orc.F16, f"one-launch block kernels n={n} nb={nb}")


def test_headline_layout_lu_in_hip_graph(handle, gpu):
    """The headline layout (fp16 storage, fp32 buffers, TC32, ib = 32, the
    one-launch block kernels, split-K updates) captured in one HIP graph:
    under capture the driver runs on the one stream and only records its
    kernels (the block kernels' sync words are cleared by a kernel inside
    the graph, no memset node). Replayed three times with eager work and
    synchronisations between the replays: every replay's factors equal the
    eager call's (one stream, no lookahead) bit for bit, and the oracle."""
    import torch
    rf = _rf()
    n, nb = 2048, 512
    a32 = orc.gen_diagdom_alpha(n, 23, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    info = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    try:
        handle.set_lookahead(False, 0)
        de = to_dev(a16, gpu)
        assert rf.lu_nopiv(handle, de, nb, algo=rf.LL, cfg=cfg, info=info) == 0
        torch.cuda.synchronize()
        assert int(info.item()) == 0
        eager = de.clone()
        handle.set_lookahead(True, 0)  # (capture itself keeps the one stream)
        buf = to_dev(a16, gpu)
        a0 = buf.clone()
        gs = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        old = handle.stream
        try:
            handle.set_stream(gs)
            with torch.cuda.graph(graph, stream=gs):
                rf.lu_nopiv(handle, buf, nb, algo=rf.LL, cfg=cfg, info=info)
        finally:
            handle.set_stream(old)
        for rep in range(3):
            buf.copy_(a0)
            info.fill_(-7)
            torch.cuda.synchronize()
            with torch.cuda.stream(gs):
                graph.replay()
            torch.cuda.synchronize()
            _ = (buf[:100].float() * 2.0).sum().item()  # eager work between replays
            assert int(info.item()) == 0, rep
            assert torch.equal(buf.view(torch.int16), eager.view(torch.int16)), rep
    finally:
        handle.set_lookahead(True, 0)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=32), a16, nb)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    _check(to_np(buf), o, f64, orc.F16, f"graph-replayed LU n={n} nb={nb}")


def test_block_kernels_wait_limit_reports_error(handle, gpu):
    """Failure handling of the one-launch block kernels, forced from the
    host: with one poll per wait the helpers give up waiting for the
    diagonal role's first slice (every kernel still runs to its end, all
    accesses in bounds), remifa_lu_nopiv reports REMIFA_ERR_HIP with its
    message, and the next factorization with the default limit runs normally
    and gives the two-launch form's bits."""
    import torch
    rf = _rf()
    n, nb = 2048, 512
    a32 = orc.gen_diagdom_alpha(n, 17, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    try:
        handle.set_block_kernels(True, 0, 1)
        da = to_dev(a16, gpu)
        with pytest.raises(rf.RemifaError) as ei:
            rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg)
        assert ei.value.status == 1000 and "timed out" in str(ei.value), str(ei.value)
        out = {}
        for one in (True, False):
            handle.set_block_kernels(one, 0, 0)
            da = to_dev(a16, gpu)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            out[one] = to_np(da)
    finally:
        handle.set_block_kernels(True, 0, 0)
    assert out[True].tobytes() == out[False].tobytes()


@pytest.mark.parametrize("n,nb,ld_pad,lookahead", [(1536, 512, 0, True), (2048, 256, 8, True),
                                                   (1536, 512, 0, False), (1280, 256, 0, True)])
def test_large_tile_updates_match_oracle(handle, gpu, n, nb, ld_pad, lookahead):
    """The headline layout (fp16 storage, fp32 buffers, TC32) with every
    outer update forced onto the 256 x 256 LDS-DMA kernel (the default
    takes it only from K = 2048, which these orders do not reach): the
    second stream's K = ofs - nb part, the K = nb part, split-K by the cost
    model, and the unsplit path without the lookahead. Against the oracle
    and against the default tiles."""
    import torch
    rf = _rf()
    a32 = orc.gen_diagdom_alpha(n, 11 + n, 1.1)
    a16 = orc.f32_to_f16_bits(a32)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=32)
    out = {}
    try:
        handle.set_lookahead(lookahead, 0)
        for tile in (4, -1):
            handle.set_gemm_tile(tile)
            da = to_dev(a16, gpu, ld=n + ld_pad)
            assert rf.lu_nopiv(handle, da, nb, algo=rf.LL, cfg=cfg) == 0
            torch.cuda.synchronize()
            out[tile] = to_np(da)
    finally:
        handle.set_gemm_tile(-1)
        handle.set_lookahead(True, 0)
    o = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=32), a16, nb)
    f64 = orc.lu_f64(orc.round_to(orc.F16, a32))
    _check(out[4], o, f64, orc.F16, f"256x256 updates n={n} nb={nb}")
    assert_parity(out[4], out[-1], f"256x256 vs default tiles n={n} nb={nb}")


@pytest.mark.parametrize("alpha", [1.01, 1.1, 2.0])
def test_c2_fp32_storage_fp16_update_alpha_sweep(handle, gpu, alpha):
    """Config 2 at n = 2048: fp32 storage with TC32 outer updates
    (lu_nopiv_ll_f32_fp32xtc32, src/lu_nopiv_ll.cxx:74-87: I = FP32, O = TC32,
    ib = 8), nb = 512, alpha in {1.01, 1.1, 2} (the diagonal dominance that
    decides how far the fp16 update can be trusted). Factors against the
    oracle; the reference harness's check solve (b = A 1, fp32 getrs with
    identity pivots) on the GPU, with nwerr / cwerr computed on the GPU and
    compared with the oracle's measures of the same solution; the factor
    error on the GPU against the oracle's."""
    import torch
    rf = _rf()
    n, nb = 2048, 512
    a32 = orc.gen_diagdom_alpha(n, 7, alpha)
    g, info = _run_gpu_helper(handle, gpu, "lu_nopiv_ll_f32_fp32xtc32", a32, nb)
    assert info == 0
    row = [r for r in rf.REFERENCE_HELPERS if r[0] == "lu_nopiv_ll_f32_fp32xtc32"][0]
    o, _ = _oracle(row, a32, nb)
    _check(g, o, orc.lu_f64(a32), orc.F32, f"C2 fp32xtc32 n={n} nb={nb} alpha={alpha}")
    da = to_dev(a32, gpu)
    dlu = to_dev(g, gpu)
    b = a32.astype(np.float64) @ np.ones(n)  # tests/testing_lu_gpu.hxx:316-351
    xb = to_dev(b.astype(np.float32), gpu)
    rf.getrs_nopiv(handle, dlu, xb)
    x = to_np(xb)[:, 0].astype(np.float64)
    m = rf.error_measures(handle, da, to_dev(x, gpu), to_dev(b, gpu), lu=dlu)
    assert m["nwerr"] == pytest.approx(orc.backward_error(a32, b, x), rel=1e-9)
    cw, cwa = orc.compwise_error(a32, b, x, g)
    assert m["cwerr"] == pytest.approx(cw, rel=1e-9)
    assert m["cwerr_ax"] == pytest.approx(cwa, rel=1e-9)
    # the solution from fp16-updated factors: backward error at the fp16 level
    assert m["nwerr"] < 2.0**-11
    fe = rf.factor_error(handle, da, dlu)
    assert fe["factor_cwerr"] == pytest.approx(orc.factor_error(a32, g), rel=1e-9)


@pytest.mark.parametrize("lu_dt,nrhs", [("f16", 1), ("f32", 1), ("f16", 3)])
def test_error_measures_match_oracle(handle, gpu, lu_dt, nrhs):
    """remifa_error_measures (nwerr, cwerr, Oettli-Prager) and
    remifa_factor_error against the oracle's restatement of tests/errors.hxx
    on the same data (n = 333: ragged chunks and tiles)."""
    rf = _rf()
    n = 333
    a32 = orc.gen_diagdom_alpha(n, 5, 1.1)
    if lu_dt == "f16":
        lu_raw = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32,
                                            ib=32), orc.f32_to_f16_bits(a32), 64)
    else:
        lu_raw = orc.lu_nopiv("ll", orc.cfg(), a32, 64)
    lu = _f(lu_raw)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (n, nrhs))
    x = np.asfortranarray(orc.getrs_f64(lu.astype(np.float64), b))
    da, dlu = to_dev(a32, gpu), to_dev(lu_raw, gpu)
    m = rf.error_measures(handle, da, to_dev(x, gpu), to_dev(np.asfortranarray(b), gpu), lu=dlu)
    assert m["nwerr"] == pytest.approx(orc.backward_error(a32, b, x), rel=1e-9)
    cws = [orc.compwise_error(a32, b[:, k], x[:, k], lu) for k in range(nrhs)]
    assert m["cwerr"] == pytest.approx(max(c[0] for c in cws), rel=1e-9)
    assert m["cwerr_ax"] == pytest.approx(max(c[1] for c in cws), rel=1e-9)
    m0 = rf.error_measures(handle, da, to_dev(x, gpu), to_dev(np.asfortranarray(b), gpu))
    assert np.isnan(m0["cwerr"]) and m0["nwerr"] == m["nwerr"]
    fe = rf.factor_error(handle, da, dlu)
    assert fe["factor_cwerr"] == pytest.approx(orc.factor_error(a32, lu), rel=1e-9)
    L, Uu = np.tril(lu.astype(np.float64), -1) + np.eye(n), np.triu(lu.astype(np.float64))
    e = a32.astype(np.float64) - L @ Uu
    assert fe["factor_abs"] == pytest.approx(np.abs(e).max(), rel=1e-9)
    assert fe["factor_rel_fro"] == pytest.approx(
        np.linalg.norm(e) / np.linalg.norm(a32.astype(np.float64)), rel=1e-9)
