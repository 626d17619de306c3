"""Shared helpers for the tests: numpy (Fortran order) <-> column-major torch
views, and the tolerance rules used by the parity tests."""
from __future__ import annotations

import numpy as np


def to_dev(a: np.ndarray, device, ld: int | None = None):
    """Fortran-order numpy matrix -> column-major torch view on `device`.
    uint16 arrays are fp16 bit patterns (stored as torch.float16)."""
    import torch
    a = np.asarray(a)
    if a.ndim == 1:
        a = a[:, None]
    m, n = a.shape
    ld = m if ld is None else ld
    if a.dtype == np.uint16:
        buf = torch.zeros((n, ld), dtype=torch.float16)
        buf.view(torch.int16)[:, :m] = torch.from_numpy(np.ascontiguousarray(a.T).view(np.int16))
    elif a.dtype == np.uint8:
        buf = torch.zeros((n, ld), dtype=torch.uint8)
        buf[:, :m] = torch.from_numpy(np.ascontiguousarray(a.T))
    else:
        buf = torch.zeros((n, ld), dtype={np.float32: torch.float32, np.float64: torch.float64}[a.dtype.type])
        buf[:, :m] = torch.from_numpy(np.ascontiguousarray(a.T))
    return buf.to(device).t()[:m]


def to_np(t) -> np.ndarray:
    """column-major torch view -> Fortran-order numpy (fp16 as uint16 bits)."""
    import torch
    c = t.t().contiguous().cpu()
    if c.dtype == torch.float16:
        return np.asfortranarray(c.view(torch.int16).numpy().view(np.uint16).T)
    if c.dtype in (torch.uint8, torch.float8_e4m3fn):
        return np.asfortranarray(c.view(torch.uint8).numpy().T)
    if c.dtype == torch.bfloat16:
        return np.asfortranarray(c.view(torch.int16).numpy().view(np.uint16).T)
    return np.asfortranarray(c.numpy().T)


def split_lu(f: np.ndarray):
    """(strict lower part with unit diagonal, upper part) of in-place factors."""
    f = np.asarray(f, np.float64)
    L = np.tril(f, -1) + np.eye(f.shape[0])
    U = np.triu(f)
    return L, U


def factor_distance(fa: np.ndarray, fb: np.ndarray) -> float:
    """max over L and U of max|a - b| / max|b| (each factor normalised by
    its own magnitude)."""
    La, Ua = split_lu(fa)
    Lb, Ub = split_lu(fb)
    dl = np.abs(La - Lb).max() / max(np.abs(Lb).max(), 1e-300)
    du = np.abs(Ua - Ub).max() / max(np.abs(Ub).max(), 1e-300)
    return float(max(dl, du))
