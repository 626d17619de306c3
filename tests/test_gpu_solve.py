"""GPU triangular solves, FP64 residual and iterative refinement.

getrs: the GPU solve with fp32 (or fp16) factors must agree with the FP64
substitution on the same factors within a backward-stable bound
|x - x64| <= 64 n u |x64|_inf scale (up to 32 dependent 128-row block rows
at n = 4096). The FP64 residual is checked against
the oracle's row-ordered FP64 sum within n * u_64 |A||x|. IR must reach the
normwise backward error 1e-14 (tests/errors.hxx measure, recomputed by the
oracle on the host).
"""
import numpy as np
import pytest

import oracle as orc
from helpers import to_dev, to_np

pytestmark = pytest.mark.gpu


def _rf():
    import remifa_amd as rf
    return rf


@pytest.mark.parametrize("n,nrhs,lu_t", [(2, 1, "f32"), (63, 1, "f32"), (64, 3, "f32"),
                                         (300, 16, "f32"), (513, 20, "f16"), (1000, 1, "f16"),
                                         (2500, 1, "f32"), (3000, 1, "f16"), (4096, 1, "f16")])
@pytest.mark.parametrize("trans", [False, True])
def test_getrs_matches_f64(handle, gpu, n, nrhs, lu_t, trans):
    import torch
    rf = _rf()
    a = orc.gen_diagdom_alpha(n, 3, 1.2)
    f = orc.lu_nopiv("rl", orc.cfg(), a, 64)
    if lu_t == "f16":
        fb = orc.f32_to_f16_bits(f)
        f = orc.f16_bits_to_f32(fb)
        dlu = to_dev(fb, gpu, ld=n + 2)
    else:
        dlu = to_dev(f, gpu, ld=n + 2)
    b = orc.gen_uniform(n, nrhs, 17)
    db = to_dev(b, gpu, ld=n + 1)
    rf.getrs_nopiv(handle, dlu, db, trans=trans)
    torch.cuda.synchronize()
    x = to_np(db).astype(np.float64)
    x64 = orc.getrs_f64(f.astype(np.float64), b.astype(np.float64), trans=trans)
    scale = np.abs(x64).max()
    assert np.abs(x - x64).max() <= 64 * n * 2.0**-24 * scale


@pytest.mark.parametrize("n,lu_t", [(1000, "f16"), (4096, "f16"), (1000, "f32"), (2500, "f32")])
def test_getrs_transposed_aligned_rows(handle, gpu, n, lu_t):
    """The transposed solves with 16-byte aligned rows of the stored factors
    (ld = n, a multiple of 8 / 4): each lane's row segment is read as vectors,
    the matrix edge and the last partial block as single entries. Against the
    FP64 solve, and bit for bit against the same factors at ld = n + 2, where
    every entry is read singly."""
    import torch
    rf = _rf()
    a = orc.gen_diagdom_alpha(n, 5, 1.2)
    f = orc.lu_nopiv("rl", orc.cfg(), a, 64)
    if lu_t == "f16":
        fb = orc.f32_to_f16_bits(f)
        f = orc.f16_bits_to_f32(fb)
    else:
        fb = f
    b = orc.gen_uniform(n, 1, 29)
    xs = []
    for pad in (0, 2):
        dlu = to_dev(fb, gpu, ld=n + pad)
        db = to_dev(b, gpu)
        rf.getrs_nopiv(handle, dlu, db, trans=True)
        torch.cuda.synchronize()
        xs.append(to_np(db))
    assert xs[0].tobytes() == xs[1].tobytes()
    x64 = orc.getrs_f64(f.astype(np.float64), b.astype(np.float64), trans=True)
    scale = np.abs(x64).max()
    assert np.abs(xs[0].astype(np.float64) - x64).max() <= 64 * n * 2.0**-24 * scale


def test_residual_f64(handle, gpu):
    import torch
    rf = _rf()
    n, nrhs = 700, 5
    a = orc.gen_diagdom_alpha(n, 8, 1.05)
    rng = np.random.default_rng(0)
    x = np.asfortranarray(rng.standard_normal((n, nrhs)))
    b = np.asfortranarray(rng.standard_normal((n, nrhs)))
    da, dx, db = to_dev(a, gpu), to_dev(x, gpu), to_dev(b, gpu)
    dr = to_dev(np.zeros((n, nrhs), order="F"), gpu)
    rf.residual_f64(handle, da, dx, db, dr)
    torch.cuda.synchronize()
    r = to_np(dr)
    ref = orc.residual_f64(a, x, b)
    bound = (n + 2) * 2.0**-53 * (np.abs(a.astype(np.float64)) @ np.abs(x) + np.abs(b))
    assert np.all(np.abs(r - ref) <= 2 * bound)


@pytest.mark.parametrize("n,nrhs,alpha", [(512, 1, 1.05), (512, 16, 1.05), (1000, 4, 1.2),
                                           (2048, 1, 1.05)])
def test_ir_reaches_1e14(handle, gpu, n, nrhs, alpha):
    import torch
    rf = _rf()
    a = orc.gen_diagdom_alpha(n, 21, alpha)
    a16 = orc.f32_to_f16_bits(a)
    da16 = to_dev(a16, gpu)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=8)
    assert rf.lu_nopiv(handle, da16, 128, algo=rf.LL, cfg=cfg) == 0
    b = orc.gen_uniform(n, nrhs, 5).astype(np.float64)
    da, db = to_dev(a, gpu), to_dev(np.asfortranarray(b), gpu)
    dx = to_dev(np.zeros((n, nrhs), order="F"), gpu)
    it, be, hist = rf.ir_solve(handle, da, da16, db, dx, tol=1e-14, maxit=20)
    torch.cuda.synchronize()
    assert it >= 0, hist
    x = to_np(dx)
    host_err = orc.backward_error(a, b, x)
    assert host_err <= 1e-14
    # the GPU's own measure agrees with the host's; at convergence both
    # residuals are mostly the rounding of their own n-term FP64 sums
    # (different orders), so they agree to a few units of 2^-53 in the
    # normwise measure, not to a fraction of it
    assert abs(host_err - be) <= 4e-15 + 0.1 * host_err
