"""Differentiable solve X = A^-1 B on the library's kernels.

Forward: the FP16-MFMA no-pivot LU of A (fp32 storage with TC32 updates,
src/lu_nopiv_ll.cxx:131, or the half-footprint fp16 storage with fp32 panels,
src/lu_nopiv_ll.cxx:618) followed by iterative refinement with the FP64
residual on the GPU until the normwise backward error is <= tol.

Backward (hand-written adjoint, reusing the stored factors):
  gB = A^-T gX   -- transposed triangular solves U^T, L^T with the same
                    factors, refined with the FP64 residual B - A^T X;
  gA = -gB X^T   -- one rank-nrhs update on the matrix cores (fp32).
Every computation runs in libremifa_hip.so; torch provides memory, streams
and the autograd graph.
"""
from __future__ import annotations

import torch

from . import api
from ._lib import FP32, RemifaError

_handles: dict = {}


def _handle(device: torch.device) -> api.Handle:
    key = (device.index, torch.cuda.current_stream(device).cuda_stream)
    h = _handles.get(key)
    if h is None:
        h = api.Handle(torch.cuda.current_stream(device))
        _handles[key] = h
    return h


class LUSolve(torch.autograd.Function):
    """X = A^-1 B with A square fp32 (no pivoting: A must admit an LU
    factorization without row exchanges, e.g. diagonally dominant)."""

    @staticmethod
    def forward(ctx, A: torch.Tensor, B: torch.Tensor, nb: int = 512, storage: str = "f32",
                tol: float = 1e-14, maxit: int = 30):
        if A.dim() != 2 or A.shape[0] != A.shape[1]:
            raise ValueError("A must be square")
        if A.dtype != torch.float32:
            raise TypeError("A must be float32 (the residual is computed from it in FP64)")
        if not A.is_cuda or B.device != A.device:
            raise ValueError("A and B must be on the same GPU")
        vec = B.dim() == 1
        B2 = B.unsqueeze(1) if vec else B
        if B2.shape[0] != A.shape[0]:
            raise ValueError("B must have as many rows as A")
        h = _handle(A.device)
        n = A.shape[0]
        a = api.to_colmajor(A)  # fp32 A for the residual (kept for backward)
        info = torch.zeros(1, dtype=torch.int32, device=A.device)
        if storage == "f16":
            lu = api.to_colmajor(A, dtype=torch.float16)
            api.lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32(h, n, n, nb, lu, n, info)
        elif storage == "f32":
            lu = api.to_colmajor(A)
            cpy = api.colmajor_empty(n, n, dtype=torch.float16, device=A.device)
            api.lu_nopiv_ll_f32_tc32xtc32(h, n, n, nb, lu, n, info, cpy, n)
        else:
            raise ValueError("storage must be 'f32' or 'f16'")
        bad = int(info.item())
        if bad:
            raise RemifaError(1002, f"LUSolve: zero or non-finite pivot at column {bad}")
        b = api.to_colmajor(B2, dtype=torch.float64)
        x = api.colmajor_empty(n, B2.shape[1], dtype=torch.float64, device=A.device)
        it, err, _ = api.ir_solve(h, a, lu, b, x, tol=tol, maxit=maxit)
        if it < 0:
            raise RemifaError(1002, f"LUSolve: refinement stopped at backward error {err:.3e}")
        ctx.save_for_backward(a, lu, x)
        ctx.vec, ctx.tol, ctx.maxit = vec, tol, maxit
        ctx.b_dtype = B.dtype
        ctx.iters, ctx.bwd_err = it, err
        out = x.to(B.dtype).contiguous()
        return out.squeeze(1) if vec else out

    @staticmethod
    def backward(ctx, gX: torch.Tensor):
        a, lu, x = ctx.saved_tensors
        h = _handle(a.device)
        n = a.shape[0]
        g = gX.unsqueeze(1) if ctx.vec else gX
        gb = api.to_colmajor(g, dtype=torch.float64)
        y = api.colmajor_empty(n, gb.shape[1], dtype=torch.float64, device=a.device)
        # gB = A^-T gX (transposed solves with the stored factors, FP64 residual)
        it, err, _ = api.ir_solve(h, a, lu, gb, y, tol=ctx.tol, maxit=ctx.maxit, trans=True)
        if it < 0:
            raise RemifaError(1002, f"LUSolve backward: refinement stopped at {err:.3e}")
        gA = None
        if ctx.needs_input_grad[0]:
            # gA = -gB X^T: rank-nrhs update on the matrix cores (fp32)
            gbf = api.to_colmajor(y, dtype=torch.float32)
            xt = api.to_colmajor(x.t(), dtype=torch.float32)
            gA = api.colmajor_empty(n, n, dtype=torch.float32, device=a.device)
            api.gemmex(h, FP32, gbf, xt, gA, alpha=-1.0, beta=0.0)
            gA = gA.contiguous()
        gB = None
        if ctx.needs_input_grad[1]:
            gB = y.to(ctx.b_dtype).contiguous()
            if ctx.vec:
                gB = gB.squeeze(1)
        return gA, gB, None, None, None, None


def solve(A: torch.Tensor, B: torch.Tensor, nb: int = 512, storage: str = "f32",
          tol: float = 1e-14, maxit: int = 30) -> torch.Tensor:
    """X = A^-1 B (differentiable in A and B)."""
    return LUSolve.apply(A, B, nb, storage, tol, maxit)
