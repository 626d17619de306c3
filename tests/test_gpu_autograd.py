"""Differentiable X = A^-1 B (remifa_amd.autograd): forward against an FP64
solve, the hand-written adjoint against FP64 central finite differences and
against torch.linalg.solve's FP64 autograd gradient (the reference gradient),
and the transposed refinement it relies on."""
import numpy as np
import pytest

import oracle as orc
from helpers import to_dev, to_np

pytestmark = pytest.mark.gpu


def _problem(n, k, seed, alpha=1.1):
    a = orc.gen_diagdom_alpha(n, seed, alpha)
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, k))
    w = rng.uniform(-1, 1, (n, k))
    return a, b, w


def test_ir_solve_transposed(handle, gpu):
    import torch
    import remifa_amd as rf
    n, k = 300, 4
    a, b, _ = _problem(n, k, 5)
    da = to_dev(a, gpu)
    lu = rf.to_colmajor(da)
    cpy = rf.colmajor_empty(n, n, dtype=torch.float16, device=gpu)
    info = torch.zeros(1, dtype=torch.int32, device=gpu)
    rf.lu_nopiv_ll_f32_tc32xtc32(handle, n, n, 64, lu, n, info, cpy, n)
    db = to_dev(b, gpu)
    x = rf.colmajor_empty(n, k, dtype=torch.float64, device=gpu)
    it, err, hist = rf.ir_solve(handle, da, lu, db, x, tol=1e-14, trans=True)
    assert it >= 0 and err <= 1e-14
    want = np.linalg.solve(a.astype(np.float64).T, b)
    assert np.max(np.abs(to_np(x) - want)) <= 1e-12 * np.max(np.abs(want))
    # the refinement measure is the normwise backward error of A^T x = b
    r = b - a.astype(np.float64).T @ to_np(x)
    nw = np.max(np.abs(r), axis=0) / (np.max(np.abs(a).sum(axis=1)) * np.max(np.abs(to_np(x)), axis=0)
                                      + np.max(np.abs(b), axis=0))
    assert np.max(nw) <= 1e-14


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_forward_matches_fp64_solve(gpu, storage):
    import torch
    from remifa_amd import autograd as ag
    n, k = 256, 3
    a, b, _ = _problem(n, k, 11)
    A = torch.from_numpy(a).to(gpu)
    B = torch.from_numpy(b).to(gpu)
    X = ag.solve(A, B, nb=64, storage=storage)
    want = np.linalg.solve(a.astype(np.float64), b)
    assert X.dtype == torch.float64
    assert np.max(np.abs(X.cpu().numpy() - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_adjoint_vs_reference_gradient(gpu, storage):
    """gA, gB against torch.linalg.solve's FP64 autograd on the CPU."""
    import torch
    from remifa_amd import autograd as ag
    n, k = 200, 5
    a, b, w = _problem(n, k, 3)
    A = torch.from_numpy(a).to(gpu).requires_grad_()
    B = torch.from_numpy(b).to(gpu).requires_grad_()
    W = torch.from_numpy(w).to(gpu)
    (ag.solve(A, B, nb=64, storage=storage) * W).sum().backward()
    Ar = torch.from_numpy(a.astype(np.float64)).requires_grad_()
    Br = torch.from_numpy(b).requires_grad_()
    (torch.linalg.solve(Ar, Br) * torch.from_numpy(w)).sum().backward()
    gB, gBr = B.grad.cpu().numpy(), Br.grad.numpy()
    assert np.max(np.abs(gB - gBr)) <= 1e-12 * np.max(np.abs(gBr))
    gA, gAr = A.grad.cpu().numpy().astype(np.float64), Ar.grad.numpy()
    # gA = -gB X^T is formed in fp32 on the matrix cores (K = k products)
    assert np.max(np.abs(gA - gAr)) <= 4 * k * 2.0**-24 * np.max(np.abs(gAr)) + 1e-30


def test_adjoint_vs_fp64_finite_differences(gpu):
    """d/dA_ij and d/dB_ij of f = <W, A^-1 B> by FP64 central differences."""
    import torch
    from remifa_amd import autograd as ag
    n, k = 64, 2
    a, b, w = _problem(n, k, 21)
    A = torch.from_numpy(a).to(gpu).requires_grad_()
    B = torch.from_numpy(b).to(gpu).requires_grad_()
    (ag.solve(A, B, nb=32) * torch.from_numpy(w).to(gpu)).sum().backward()
    a64 = a.astype(np.float64)

    def f(am, bm):
        return float(np.sum(w * np.linalg.solve(am, bm)))

    rng = np.random.default_rng(0)
    h = 1e-6
    for _ in range(6):
        i, j = rng.integers(0, n, 2)
        ap, am = a64.copy(), a64.copy()
        ap[i, j] += h
        am[i, j] -= h
        fd = (f(ap, b) - f(am, b)) / (2 * h)
        assert abs(A.grad[i, j].item() - fd) <= 1e-5 * max(abs(fd), 1e-3)
        c = rng.integers(0, k)
        bp, bm = b.copy(), b.copy()
        bp[i, c] += h
        bm[i, c] -= h
        fd = (f(a64, bp) - f(a64, bm)) / (2 * h)
        assert abs(B.grad[i, c].item() - fd) <= 1e-7 * max(abs(fd), 1e-3)


def test_vector_rhs_and_float32_b(gpu):
    import torch
    from remifa_amd import autograd as ag
    n = 128
    a, b, _ = _problem(n, 1, 8)
    A = torch.from_numpy(a).to(gpu).requires_grad_()
    B = torch.from_numpy(b[:, 0].astype(np.float32)).to(gpu).requires_grad_()
    X = ag.solve(A, B, nb=64)
    assert X.shape == (n,) and X.dtype == torch.float32
    X.sum().backward()
    assert B.grad.shape == (n,) and A.grad.shape == (n, n)
    want = np.linalg.solve(a.astype(np.float64).T, np.ones(n))
    assert np.max(np.abs(B.grad.cpu().numpy() - want)) <= 2.0**-22 * np.max(np.abs(want))
