"""The harness rows on the GPU: the other generators (gen_randmat,
gen_arrowhead) bit-exact against the oracle's restatements, equilibration
(LAPACK sgeequb / slaqge semantics) bit-exact against the oracle, the
conversion check (--check-conv) against numpy, and the runtime-dispatched
test driver tests/testing_lu_gpu.py over the reference's option matrix,
including the rocSOLVER no-pivot comparator (--algo=cusolver) and the
condition-number generator (--cond)."""
import json
import os

import numpy as np
import pytest

import oracle as orc
from helpers import to_dev, to_np

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("m,n,seed", [(37, 29, 5), (256, 256, 1), (1000, 300, 1), (4097, 130, 9)])
def test_gen_randmat_matches_oracle(handle, gpu, m, n, seed):
    import remifa_amd as rf
    a = rf.colmajor_empty(m, n, device=gpu, ld=m + 3)
    rf.gen_randmat(handle, a, seed=seed)
    assert np.array_equal(to_np(a).view(np.uint32), orc.gen_randmat(m, n, seed).view(np.uint32))


@pytest.mark.parametrize("n,lo,hi", [(100, 0.0, 1.0), (300, 0.0, 1.0), (300, -1.0, 1.0),
                                     (600, -1.0, 1.0), (1030, 0.0, 1.0)])
def test_gen_arrowhead_matches_oracle(handle, gpu, n, lo, hi):
    import remifa_amd as rf
    a = rf.colmajor_empty(n, n, device=gpu, ld=n + 1)
    rf.gen_arrowhead(handle, a, lo, hi)
    assert np.array_equal(to_np(a).view(np.uint32), orc.gen_arrowhead(n, lo, hi).view(np.uint32))


def _geequb_gpu(handle, gpu, a):
    import torch
    import remifa_amd as rf
    m, n = a.shape
    da = to_dev(a, gpu, ld=m + 2)
    r = torch.zeros(m, dtype=torch.float32, device=gpu)
    c = torch.zeros(n, dtype=torch.float32, device=gpu)
    res = rf.geequb(handle, da, r, c)
    return da, r, c, res


def test_geequb_matches_oracle_and_lapack(handle, gpu):
    """the fixture's badly scaled matrices (tests/golden/geequb.json, LAPACK's
    own results) and a gen_diagdom matrix (a_ii = n, a power of two)"""
    g = json.load(open(os.path.join(GOLDEN, "geequb.json")))
    mats = [np.array(c["a_bits"], np.uint32).view(np.float32).reshape((c["m"], c["n"]), order="F")
            for c in g["cases"]] + [orc.gen_diagdom(512)]
    for a in mats:
        _, r, c, res = _geequb_gpu(handle, gpu, a)
        ro, co, rowcnd, colcnd, amax, info = orc.geequb(a)
        assert res["info"] == info == 0
        assert np.array_equal(r.cpu().numpy().view(np.uint32), ro.view(np.uint32))
        assert np.array_equal(c.cpu().numpy().view(np.uint32), co.view(np.uint32))
        assert (res["rowcnd"], res["colcnd"], res["amax"]) == (rowcnd, colcnd, amax)


def test_geequb_reports_zero_row_and_column(handle, gpu):
    a = np.asfortranarray(np.eye(40, dtype=np.float32))
    a[17, 17] = 0.0
    assert _geequb_gpu(handle, gpu, a)[3]["info"] == 18
    b = np.asfortranarray(np.ones((30, 20), np.float32))
    b[:, 7] = 0.0
    assert _geequb_gpu(handle, gpu, b)[3]["info"] == 30 + 8


@pytest.mark.parametrize("case", ["N", "R", "C", "B"])
def test_laqge_matches_oracle(handle, gpu, case):
    import remifa_amd as rf
    rng = np.random.default_rng(4)
    m, n = 300, 260
    a = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    if case in ("R", "B"):
        a *= (10.0 ** rng.integers(-5, 5, (m, 1))).astype(np.float32)
    if case in ("C", "B"):
        a *= (10.0 ** rng.integers(-5, 5, (1, n))).astype(np.float32)
    a = np.asfortranarray(a)
    da, r, c, res = _geequb_gpu(handle, gpu, a)
    eq = rf.laqge(handle, da, r, c, res["rowcnd"], res["colcnd"], res["amax"])
    ro, co, rowcnd, colcnd, amax, _ = orc.geequb(a)
    want, eq_o = orc.laqge(a, ro, co, rowcnd, colcnd, amax)
    assert eq == eq_o == case
    assert np.array_equal(to_np(da).view(np.uint32), want.view(np.uint32))


def test_conv_error_matches_numpy(handle, gpu):
    import torch
    import remifa_amd as rf
    a = orc.gen_diagdom_alpha(500, 3, 1.1)
    a[3, 4] = 0.0
    da = to_dev(a, gpu)
    d16 = rf.colmajor_empty(500, 500, dtype=torch.float16, device=gpu)
    rf.convert(handle, da, d16)
    ce = rf.conv_error(handle, da, d16)
    t = a.astype(np.float16).astype(np.float64)
    e = np.abs(a.astype(np.float64) - t)
    rel = np.where(a != 0, e / np.abs(a.astype(np.float64)), e)
    assert ce["abs"] == e.max() and ce["rel"] == rel.max()
    assert 0 < ce["rel"] <= 2.0**-11


def _cli(args):
    from testing_lu_gpu import run
    return run(args)


@pytest.mark.parametrize("args,tol", [
    (["--algo=remifa-ll", "--fp16", "--in-upd=tc32", "--out-upd=tc32", "--out-pnl=fp32",
      "--in-pnl=fp32", "--m", "1024", "--nb", "256", "--ib", "32", "--diagdom", "--check-conv"],
     2.0**-10),
    (["--algo=remifa-ll", "--fp32", "--in-upd=fp32", "--out-upd=tc32", "--m", "768", "--nb", "256",
      "--alpha", "1.01", "--seed", "3"], 2.0**-10),
    (["--algo=remifa-rl", "--fp32", "--m", "512", "--nb", "128", "--cond", "2"], 1e-5),
    (["--algo=remifa-ll", "--fp32", "--in-upd=tc32", "--out-upd=tc32", "--m", "600", "--nb",
      "128", "--arrowhead-pos-neg", "--scale"], 2.0**-10),
    (["--algo=remifa-ll", "--fp16", "--in-upd=fp16", "--out-upd=fp16", "--m", "300", "--nb",
      "64", "--arrowhead"], 2.0**-7),
    (["--algo=cusolver", "--fp32", "--m", "1024", "--diagdom"], 1e-6),
    (["--algo=remifa-ll", "--fp32", "--nb", "4", "--mtx",
      os.path.join(GOLDEN, "small_sym.mtx")], 1e-6),
])
def test_test_driver_runs_the_option_matrix(gpu, args, tol):
    """the driver factors, solves A x = A 1 and reports the reference's
    nwerr / cwerr (GPU, FP64) at the level the precision tuple allows"""
    out = _cli(args + ["--json"])
    assert out["info"] == 0
    assert out["nwerr"] < tol and out["cwerr"] < 4 * tol, out
    assert out["gflops"] > 0
    if "--check-conv" in args:
        assert 0 < out["conv_cwerr"] <= 2.0**-11


def test_test_driver_cond_generator_has_the_condition_number(gpu):
    """--cond C: A = Q D Q^T with D = diag(10^(-C i/(n-1))): symmetric, with
    eigenvalues D (fp32 construction) and 2-norm condition number 10^C"""
    import torch
    import remifa_amd as rf
    from testing_lu_gpu import gen_posdef_cond
    h = rf.Handle()
    n = 256
    a = gen_posdef_cond(h, n, 3.0)
    torch.cuda.synchronize()
    a64 = to_np(a).astype(np.float64)
    assert np.abs(a64 - a64.T).max() <= 1e-6
    w = np.linalg.eigvalsh((a64 + a64.T) / 2)
    want = np.sort(10.0 ** (-3.0 * np.arange(n) / (n - 1)))
    assert np.allclose(w, want, rtol=0, atol=1e-5)
    assert np.linalg.cond(a64) == pytest.approx(1e3, rel=0.05)


def test_test_driver_rejects_what_is_not_built(gpu):
    for args in (["--latmr", "1"], ["--fp64"], ["--ngpu", "2"]):
        with pytest.raises(SystemExit):
            _cli(args)
