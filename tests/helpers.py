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


# ---------------------------------------------------------------- entrywise parity
# GPU and oracle run the same operations with the same rounding points and
# differ only in the summation order inside the matrix-core GEMMs. The rule
# measures each entry's difference against the size of the terms that make
# it: r_ij = |g_ij - o_ij| / (u_W s_ij) with s = |L||U| of the oracle's
# factors (divided by |u_jj| below the diagonal, where l_ij = (...) / u_jj)
# and u_W the unit roundoff of the storage format. Two results that differ
# only in summation order (the oracle against itself with k-descending sums,
# oracle.set_sum_order(1)) give, in fp16 storage, >= 99.99% bit-identical
# entries and r_99.9% = 0; in fp32 storage r_99.9% <= 91 (an fp16 operand copy
# that rounds the other way spreads into later entries). A misplaced rounding
# point (fp16 where fp32 buffers are kept, a dropped or extra operand
# rounding, bf16 for fp16 operands) gives, in fp16 storage, <= 99.0%
# identical entries and r_99.9% >= 1.1, in fp32 storage r_99.9% >= 712
# (tests/test_oracle.py plants such mistakes in the oracle and checks that
# this rule rejects them).
F16_MIN_EQUAL = 0.999  # fp16 storage: fraction of bit-identical entries
F16_MAX_R999 = 0.5     # fp16 storage: 99.9th percentile of r
F16_MAX_R = 8.0        # fp16 storage: largest r (two ulps at the entry's own size is r = 4)
F32_MIN_EQUAL = 0.5    # fp32 storage: fraction of bit-identical entries
F32_MAX_R999 = 256.0   # fp32 storage: 99.9th percentile of r


def _ordinal16(bits: np.ndarray) -> np.ndarray:
    x = np.asarray(bits).astype(np.int32)
    return np.where(x & 0x8000, -(x & 0x7FFF), x & 0x7FFF)


def _ordinal32(vals: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(vals, np.float32)).view(np.int32).astype(np.int64)
    return np.where(x < 0, -(x & 0x7FFFFFFF), x)


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Entrywise distance in units in the last place: a, b are fp16 bit
    patterns (uint16) or fp32 arrays. Adjacent representable values are 1
    apart, +0 and -0 are 0 apart."""
    if np.asarray(a).dtype == np.uint16:
        return np.abs(_ordinal16(a) - _ordinal16(b))
    return np.abs(_ordinal32(a) - _ordinal32(b))


def _values(raw: np.ndarray) -> np.ndarray:
    raw = np.asarray(raw)
    if raw.dtype == np.uint16:
        import oracle as orc
        return orc.f16_bits_to_f32(raw).astype(np.float64)
    return raw.astype(np.float64)


def entry_scale(o_raw: np.ndarray) -> np.ndarray:
    """s = |L||U| of in-place factors, divided by |u_jj| strictly below the
    diagonal (the magnitude of the terms each entry is computed from)"""
    o = _values(o_raw)
    n = o.shape[0]
    L = np.tril(o, -1) + np.eye(n)
    U = np.triu(o)
    S = np.abs(L) @ np.abs(U)
    d = np.abs(np.diag(U)).copy()
    d[d == 0] = 1.0
    return np.where(np.tril(np.ones((n, n), bool), -1), S / d[None, :], S)


def parity_stats(g: np.ndarray, o: np.ndarray) -> dict:
    d = ulp_distance(g, o)
    u = 2.0**-11 if np.asarray(g).dtype == np.uint16 else 2.0**-24
    diff = np.abs(_values(g) - _values(o))
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(diff > 0, diff / (u * entry_scale(o)), 0.0)
    return {"frac_equal": float((d == 0).mean()), "max_ulp": int(d.max()),
            "r999": float(np.quantile(r, 0.999)), "rmax": float(r.max()), "entries": int(d.size)}


def parity_ok(g: np.ndarray, o: np.ndarray) -> tuple[bool, dict]:
    """The entrywise rule for factors in storage format (fp16 bits or fp32)"""
    s = parity_stats(g, o)
    if np.asarray(g).dtype == np.uint16:
        ok = (s["frac_equal"] >= F16_MIN_EQUAL and s["r999"] <= F16_MAX_R999
              and s["rmax"] <= F16_MAX_R)
    else:
        ok = s["frac_equal"] >= F32_MIN_EQUAL and s["r999"] <= F32_MAX_R999
    return ok, s


def log_parity(label: str, stats: dict) -> None:
    """Append the statistics to $REMIFA_PARITY_LOG (JSON lines) if set."""
    import json
    import os
    path = os.environ.get("REMIFA_PARITY_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"case": label, **stats}) + "\n")


def assert_parity(g: np.ndarray, o: np.ndarray, label: str = "") -> dict:
    ok, s = parity_ok(g, o)
    log_parity(label, s)
    assert ok, f"{label}: GPU vs oracle {s}"
    return s
