"""The headline workload at its full size, checked on the host.

n = 32768, fp16 storage, fp32 inner buffers, TC32 outer updates, nb = 512,
ib = 32, lookahead on (src/lu_nopiv_ll.cxx:618), seeded alpha = 1.05
generator, then FP64-residual refinement to 1e-14. The oracle cannot factor
this size in test time, so the checks are size-independent properties
computed from what the GPU returns:
 * info == 0 and the refinement converges;
 * the normwise and componentwise backward errors of the refined solution
   (tests/errors.hxx:22-249) recomputed on the host in FP64 from A, the
   factors and x read back (column chunks, numpy) are <= 1e-14;
 * for the unrefined solution x0 = (LU)^-1 b (fp32 solves), whose residual is
   far above FP64 rounding, the GPU's error kernels agree with the host's to
   1e-9 relative;
 * the normwise factor error ||A - LU||_F / ||A||_F on the GPU is at the
   level of fp16 storage (the factors are rounded to fp16 once per entry).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NB, IB, ALPHA, SEED = 32768, 512, 32, 1.05, 2024
CH = 512  # columns per host chunk


def _host_measures(A, LU, b, x):
    """(nwerr, cwerr, cwerr_ax) of tests/errors.hxx in FP64 on the host, A
    fp32 and LU (fp16) column-major device views, b and x numpy vectors"""
    n = A.shape[0]
    r = b.copy()
    ax = np.zeros(n)
    t = np.zeros(n)
    anorm = 0.0
    ri = np.arange(n)
    ax_abs = np.abs(x)
    for j0 in range(0, n, CH):
        j1 = min(n, j0 + CH)
        blk = A.t()[j0:j1].cpu().numpy().astype(np.float64)  # columns j0..j1 as rows
        r -= x[j0:j1] @ blk
        ab = np.abs(blk)
        ax += ax_abs[j0:j1] @ ab
        anorm = max(anorm, float(ab.sum(axis=1).max()))
        lb = np.abs(LU.t()[j0:j1].cpu().numpy().astype(np.float64))
        t += ax_abs[j0:j1] @ np.where(ri[None, :] <= np.arange(j0, j1)[:, None], lb, 0.0)
    lux = t.copy()
    for j0 in range(0, n, CH):
        j1 = min(n, j0 + CH)
        lb = np.abs(LU.t()[j0:j1].cpu().numpy().astype(np.float64))
        lux += t[j0:j1] @ np.where(ri[None, :] > np.arange(j0, j1)[:, None], lb, 0.0)
    ra = np.abs(r)
    nw = ra.max() / (anorm * np.abs(x).max() + np.abs(b).max())
    with np.errstate(divide="ignore", invalid="ignore"):
        cw = np.nanmax(np.where(ax + lux > 0, ra / (ax + lux), ra))
        cwa = np.nanmax(np.where(ax > 0, ra / ax, 0.0))
    return float(nw), float(cw), float(cwa)


@pytest.mark.timeout(900)
def test_headline_lu_and_refinement_full_size(gpu):
    import torch
    import remifa_amd as rf
    print(f"full-size headline: n={N} nb={NB} ib={IB} alpha={ALPHA} seed={SEED}", flush=True)
    h = rf.Handle()
    h.set_lookahead(True, 0)
    A = rf.colmajor_empty(N, N, device=gpu)
    rf.gen_diagdom(h, A, seed=SEED, alpha=ALPHA)
    A16 = rf.colmajor_empty(N, N, dtype=torch.float16, device=gpu)
    rf.convert(h, A, A16)
    cfg = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=rf.TC32, OP=rf.R_32F, ib=IB)
    assert rf.lu_nopiv(h, A16, NB, algo=rf.LL, cfg=cfg) == 0
    bf = rf.colmajor_empty(N, 1, device=gpu)
    rf.gen_uniform(h, bf, seed=SEED + 1)
    B = torch.zeros((1, N), dtype=torch.float64, device=gpu).t()
    B.copy_(bf)
    X = torch.zeros((1, N), dtype=torch.float64, device=gpu).t()
    it, be, hist = rf.ir_solve(h, A, A16, B, X, tol=1e-14, maxit=30)
    assert it >= 0, hist
    torch.cuda.synchronize()
    b = B[:, 0].cpu().numpy()
    x = X[:, 0].cpu().numpy()
    nw, cw, cwa = _host_measures(A, A16, b, x)
    print(f"refined: it={it} gpu nwerr={be:.3e} host nwerr={nw:.3e} cwerr={cw:.3e} "
          f"oettli-prager={cwa:.3e}", flush=True)
    assert nw <= 1e-14 and cw <= 1e-14 and cwa <= 1e-14
    m = rf.error_measures(h, A, X, B, lu=A16)
    assert m["nwerr"] <= 1e-14 and m["cwerr"] <= 1e-14 and m["cwerr_ax"] <= 1e-14
    # the unrefined solution: residual well above FP64 rounding
    x0f = bf.clone()
    rf.getrs_nopiv(h, A16, x0f)
    X0 = x0f.double()
    m0 = rf.error_measures(h, A, X0, B, lu=A16)
    nw0, cw0, cwa0 = _host_measures(A, A16, b, X0[:, 0].cpu().numpy())
    print(f"x0: gpu {m0} host nwerr={nw0:.6e} cwerr={cw0:.6e} op={cwa0:.6e}", flush=True)
    assert m0["nwerr"] == pytest.approx(nw0, rel=1e-9)
    assert m0["cwerr"] == pytest.approx(cw0, rel=1e-9)
    assert m0["cwerr_ax"] == pytest.approx(cwa0, rel=1e-9)
    assert nw0 > 1e-12  # (the comparison above is not between two roundings)
    fe = rf.factor_error(h, A, A16)
    print(f"factor error: {fe}", flush=True)
    # fp16 factors of the fp16-rounded A: ||A - LU||_F / ||A||_F a few units
    # of fp16 roundoff. (The componentwise form max |A - LU| / (|L||U|) is
    # large here and is not bounded: with a_ii ~ 1.7e4 every l_ij is below
    # fp16's smallest normal, 6.1e-5, so L is stored with absolute rather
    # than relative precision.)
    assert fe["factor_rel_fro"] < 8 * 2.0**-11
    assert fe["factor_abs"] <= 16.0  # (|a_ii| ~ 1.7e4: its fp16 ulp is 16)
