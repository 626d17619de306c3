#!/usr/bin/env python3
"""Benchmark: no-pivot LU at n=32768 on one MI355X, FP16-MFMA + FP64-residual
iterative refinement vs plain FP32.

One step = the hot path on one matrix: factor the fp16 (half-footprint) copy
of A with MFMA updates (fp16 operands, fp32 accumulation), then refine the
solution of A x = b with the FP64 residual until the normwise backward error
(tests/errors.hxx:186-249) is <= 1e-14. A (fp32, seeded, alpha = 1.05) and b
are generated on the GPU and stay resident; restoring the fp16 working copy
before each step is outside the timed region, as the reference's harness
converts before starting its timer (tests/testing_lu_gpu.hxx:514, 531).

Printed: ONE JSON line. value = LU TFLOP/s of the fp16 factorization
(2n^3/3 / t_factor, tests/testing_lu_gpu.hxx:1412-1414); ms_per_step = time to
solution (factor + refinement). Beside it, in the same run:
 * the plain FP32 GPU factorization (left- and right-looking, the faster one
   is the comparator), with an FP32 solve and with the same FP64 refinement;
 * the BF16 and FP8 update variants and the nrhs = 16 refinement;
 * the reference's error measures (tests/errors.hxx) of every result on the
   GPU: normwise and componentwise backward error, Oettli-Prager form, factor
   error, and the forward error against an FP64 solution (rocSOLVER
   dgetrf_npvt + dgetrs); the headline's backward error is also recomputed
   on the host in FP64;
 * rocSOLVER's no-pivot LU (sgetrf_npvt) on the same matrix, as the
   reference times cuSOLVER getrf with a NULL pivot array
   (tests/testing_lu_gpu.hxx:390-468);
 * config 2 (fp32 storage, TC32 updates, n = 16384, alpha in {1.01, 1.1, 2}),
   config 4 (20000 batched LUs + solves in one HIP graph) and config 5 (the
   differentiable solve, A 8192^2, B 8192 x 64) as sub-records;
 * cpu_baseline: the oracle's CPU restatement of lu_nopiv_ll_f32 on a
   bounded sample of the headline generator.
The roofline figure of the dominant kernel (the MFMA update GEMM) comes from
hipEvents recorded around every GEMM launch on the stream it runs on, in an
instrumented pass after the timed steps. The process exits non-zero (after
printing the line) if a factorization reports a zero pivot, a refinement does
not converge or the host recheck exceeds 1e-14.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LU TFLOP/s & time to 1e-14 solve at n=32k: FP16-MFMA+IR vs plain FP32"
PEAK_FP16_TFLOPS = 2500.0  # MI355X dense fp16/bf16 matrix peak
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

FAILURES: list = []


def fail(msg: str) -> None:
    FAILURES.append(msg)
    print(f"bench: FAILED: {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=32768)
    p.add_argument("--nb", type=int, default=512)
    p.add_argument("--ib", type=int, default=32)
    p.add_argument("--algo", choices=["ll", "rl"], default="ll")
    p.add_argument("--alpha", type=float, default=1.05)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--nrhs", type=int, default=1)
    p.add_argument("--fp32-steps", type=int, default=1)
    p.add_argument("--fp32-nb", type=int, default=0)
    p.add_argument("--cpu-n", type=int, default=0, help="CPU baseline sample size (0 = auto)")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--pmc-calibrate", action="store_true",
                   help="first copy 256 MiB device to device (a dispatch of known HBM "
                        "traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    p.add_argument("--update", choices=["fp16", "bf16", "fp8"], default="fp16",
                   help="outer-update operands of the timed factorization")
    p.add_argument("--no-lookahead", action="store_true",
                   help="LL: factor each outer block after its whole update (one stream)")
    p.add_argument("--reserve-cus", type=int, default=0,
                   help="LL lookahead: CUs masked off the second stream")
    p.add_argument("--no-variants", action="store_true",
                   help="skip the BF16 / FP8 update variants and the nrhs=16 refinement")
    p.add_argument("--no-errors", action="store_true",
                   help="skip the error measures (GPU factor errors, host recheck)")
    p.add_argument("--no-vendor", action="store_true", help="skip the rocSOLVER comparator")
    p.add_argument("--no-configs", action="store_true",
                   help="skip the config 2 / 4 / 5 sub-records")
    p.add_argument("--workload", choices=["lu", "batched", "autograd", "c2"], default="lu",
                   help="lu: the headline n=32768 LU + IR with every sub-record (default); "
                        "batched: 20000 LUs + solves of orders U[16,512]; autograd: X = A^-1 B "
                        "forward + backward; c2: fp32 storage with TC32 updates, n=16384")
    p.add_argument("--count", type=int, default=20000, help="batched: number of matrices")
    p.add_argument("--per-matrix-sample", type=int, default=400,
                   help="batched: matrices timed with one driver call each")
    return p.parse_args()


def ev():
    import torch
    return torch.cuda.Event(enable_timing=True)


# ---------------------------------------------------------------- measures
def host_nwerr(A, b, x, chunk: int = 512) -> float:
    """||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) (tests/errors.hxx:
    186-249) in FP64 on the host, A a column-major fp32 device view read back
    in column chunks, b and x numpy vectors"""
    import numpy as np
    n = A.shape[0]
    r = b.astype(np.float64).copy()
    anorm = 0.0
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        blk = A.t()[j0:j1].cpu().numpy().astype(np.float64)  # columns as rows
        r -= x[j0:j1] @ blk
        anorm = max(anorm, float(np.abs(blk).sum(axis=1).max()))
    return float(np.abs(r).max() / (anorm * np.abs(x).max() + np.abs(b).max()))


def measures(h, A, X, B, lu, factor: bool = True, xref=None) -> dict:
    """the reference's error measures of the solution X (FP64 device) and
    the factors lu, on the GPU; forward error against xref if given"""
    import remifa_amd as rf
    m = rf.error_measures(h, A, X, B, lu=lu)
    out = {"nwerr": m["nwerr"], "cwerr": m["cwerr"], "cwerr_oettli_prager": m["cwerr_ax"]}
    if factor:
        fe = rf.factor_error(h, A, lu)
        out["factor_error"] = fe["factor_cwerr"]
        out["factor_error_rel_fro"] = fe["factor_rel_fro"]
    if xref is not None:
        out["forward_error"] = float((X - xref).abs().max() / xref.abs().max())
    return out


# ---------------------------------------------------------------- rocSOLVER
class RocSolver:
    """ctypes binding of rocSOLVER's no-pivot LU (the reference's cuSOLVER
    getrf with devIpiv = NULL, src/wrappers.cxx:564-576) and FP64 getrs, the
    libraries torch ships, on torch's current stream"""

    def __init__(self):
        import ctypes as C
        import torch
        lib = os.path.join(os.path.dirname(torch.__file__), "lib")
        self.C = C
        self.rb = C.CDLL(os.path.join(lib, "librocblas.so"))
        self.rs = C.CDLL(os.path.join(lib, "librocsolver.so"))
        self.h = C.c_void_p()
        st = self.rb.rocblas_create_handle(C.byref(self.h))
        if st != 0:
            raise RuntimeError(f"rocblas_create_handle: {st}")
        self.rb.rocblas_set_stream(self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        for f in ("rocsolver_sgetrf_npvt", "rocsolver_dgetrf_npvt"):
            getattr(self.rs, f).argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                            C.c_void_p]
        self.rs.rocsolver_dgetrs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        self.rs.rocsolver_sgeqrf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                             C.c_void_p]
        self.rs.rocsolver_sormqr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_int]

    def getrf_npvt(self, a, info) -> None:
        n = a.shape[0]
        f = self.rs.rocsolver_sgetrf_npvt if a.dtype.itemsize == 4 else self.rs.rocsolver_dgetrf_npvt
        st = f(self.h, n, n, self.C.c_void_p(a.data_ptr()), a.stride(1),
               self.C.c_void_p(info.data_ptr()))
        if st != 0:
            raise RuntimeError(f"rocsolver getrf_npvt: status {st}")

    def dgetrs_nopiv(self, lu, ipiv, b) -> None:
        n = lu.shape[0]
        st = self.rs.rocsolver_dgetrs(self.h, 111, n, b.shape[1], self.C.c_void_p(lu.data_ptr()),
                                      lu.stride(1), self.C.c_void_p(ipiv.data_ptr()),
                                      self.C.c_void_p(b.data_ptr()), max(b.stride(1), n))
        if st != 0:
            raise RuntimeError(f"rocsolver dgetrs: status {st}")

    def sgeqrf(self, a, tau) -> None:
        """Householder QR of the fp32 column-major a in place (tau: n)"""
        m, n = a.shape
        st = self.rs.rocsolver_sgeqrf(self.h, m, n, self.C.c_void_p(a.data_ptr()), a.stride(1),
                                      self.C.c_void_p(tau.data_ptr()))
        if st != 0:
            raise RuntimeError(f"rocsolver sgeqrf: status {st}")

    def sormqr(self, left: bool, trans: bool, qr, tau, c) -> None:
        """c <- op(Q) c (left) or c op(Q) (right), Q from sgeqrf"""
        m, n = c.shape
        k = qr.shape[1]
        st = self.rs.rocsolver_sormqr(self.h, 141 if left else 142, 112 if trans else 111, m, n,
                                      k, self.C.c_void_p(qr.data_ptr()), qr.stride(1),
                                      self.C.c_void_p(tau.data_ptr()),
                                      self.C.c_void_p(c.data_ptr()), c.stride(1))
        if st != 0:
            raise RuntimeError(f"rocsolver sormqr: status {st}")

    def close(self):
        self.rb.rocblas_destroy_handle(self.h)


def fp64_reference_solution(A, B):
    """x* of A x = B in FP64 (rocSOLVER dgetrf_npvt + dgetrs with identity
    pivots): the reference solution of the forward errors"""
    import torch
    import remifa_amd as rf
    n = A.shape[0]
    rs = RocSolver()
    A64 = rf.colmajor_empty(n, n, dtype=torch.float64, device=A.device)
    A64.copy_(A)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    rs.getrf_npvt(A64, info)
    ipiv = torch.arange(1, n + 1, dtype=torch.int32, device=A.device)
    xr = rf.colmajor_empty(n, B.shape[1], dtype=torch.float64, device=A.device)
    xr.copy_(B)
    rs.dgetrs_nopiv(A64, ipiv, xr)
    torch.cuda.synchronize()
    rs.close()
    if int(info.item()) != 0:
        fail(f"rocsolver dgetrf_npvt info {int(info.item())}")
    del A64
    return xr


def bench_vendor(h, A, B, xref, nb32_ms):
    """rocSOLVER sgetrf_npvt on the same fp32 matrix: time, errors"""
    import torch
    import remifa_amd as rf
    n = A.shape[0]
    rs = RocSolver()
    A32 = rf.colmajor_empty(n, n, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    ts = []
    for k in range(2):
        A32.copy_(A)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        rs.getrf_npvt(A32, info)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    rs.close()
    out = {"routine": "rocsolver_sgetrf_npvt (torch's bundled rocSOLVER)",
           "factor_ms": round(ts[-1], 3),
           "tflops": round(2.0 * n**3 / 3.0 / (ts[-1] * 1e-3) / 1e12, 2), "info": int(info.item())}
    if out["info"] != 0:
        out["note"] = "zero pivot reported"
        return out
    xb = rf.colmajor_empty(n, B.shape[1], device=A.device)
    xb.copy_(B)
    rf.getrs_nopiv(h, A32, xb)
    X = xb.double()
    out.update(measures(h, A, X, B, A32, xref=xref))
    if nb32_ms:
        out["library_fp32_factor_speedup"] = round(ts[-1] / nb32_ms, 3)
    del A32
    return out


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(target_s: float, n_fixed: int, seed: int, alpha: float):
    """Oracle restatement of lu_nopiv_ll_f32 (nb=256, ib=8) on the host
    cores, on the headline generator (same seed and alpha) at a smaller
    order: the order is grown until one factorization takes a few seconds,
    then timed once."""
    import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    orc.set_threads(threads)
    c = orc.cfg()
    n = n_fixed or 1024
    while True:
        a = orc.gen_diagdom_alpha(n, seed, alpha)
        t0 = time.perf_counter()
        orc.lu_nopiv("ll", c, a, 256)
        dt = time.perf_counter() - t0
        if n_fixed or dt > target_s / 8 or n >= 16384:
            break
        n *= 2
    tflops = (2.0 * n**3 / 3.0) / dt / 1e12
    return {"value": tflops, "unit": "TFLOP/s", "cores": orc.get_threads(), "kind": "port",
            "sample": f"oracle LL restatement of lu_nopiv_ll_f32 (nb=256, ib=8) on the "
                      f"headline generator (seed {seed}, alpha {alpha}) at n={n}, one "
                      f"factorization, {dt:.2f} s",
            "gemm": "the oracle's own k-ordered fp32 fma loops over OpenMP threads "
                    "(no host BLAS)"}


# ---------------------------------------------------------------- config 4
def bench_batched(a):
    """20000 independent no-pivot LUs + solves, orders uniform in [16, 512],
    alpha = 1.2, fp32 storage with TC32 (fp16 matrix-core) updates: one
    persistent launch each for the factorizations and the solves, captured
    in a HIP graph, against one library call per matrix."""
    import numpy as np
    import torch

    import remifa_amd as rf

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(a.seed)
    sizes = rng.integers(16, 513, a.count)
    order = np.argsort(-sizes, kind="stable")  # largest first: LPT queue
    sizes = sizes[order]
    lds = sizes.astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lds * sizes)[:-1]])
    total = int((lds * sizes).sum())
    h = rf.Handle()
    a0 = torch.empty(total, dtype=torch.float32, device=dev)
    for i, (n, off) in enumerate(zip(sizes, offs)):
        view = a0[off:off + n * n].view(int(n), int(n)).t()
        rf.gen_diagdom(h, view, seed=a.seed + i, alpha=1.2)
    b0 = torch.empty(int(sizes.sum()), dtype=torch.float32, device=dev)
    boffs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    rf.gen_uniform(h, b0.view(1, -1).t(), seed=a.seed + 1)
    work = torch.empty_like(a0)
    bw = torch.empty_like(b0)
    ptrs = torch.tensor(work.data_ptr() + 4 * offs, dtype=torch.int64, device=dev)
    bptrs = torch.tensor(bw.data_ptr() + 4 * boffs, dtype=torch.int64, device=dev)
    n_t = torch.tensor(sizes, dtype=torch.int32, device=dev)
    ld_t = torch.tensor(lds, dtype=torch.int32, device=dev)
    info = torch.zeros(a.count, dtype=torch.int32, device=dev)
    flops = float((2.0 * sizes.astype(np.float64) ** 3 / 3.0).sum())
    bytes_ = float((8.0 * sizes.astype(np.float64) ** 2).sum())

    # size buckets (the batch is sorted largest first): one persistent
    # launch per bucket, its kernel shape from the autotuned table
    buckets = []
    for lo_b, bm in ((256, 512), (128, 256), (64, 128), (0, 64)):
        idx = np.nonzero((sizes > lo_b) & (sizes <= bm))[0]
        if len(idx):
            buckets.append((int(idx[0]), int(idx[-1]) + 1, bm))

    def run():
        for i0, i1, bm in buckets:
            rf.lu_nopiv_batched(h, rf.TC32, n_t[i0:i1], ld_t[i0:i1], ptrs[i0:i1], info[i0:i1], bm)
        rf.getrs_nopiv_batched(h, n_t, ld_t, ptrs, bptrs, 512)

    work.copy_(a0); bw.copy_(b0)
    run()  # warm-up (and workspace)
    torch.cuda.synchronize()
    # HIP graph of the two launches on the handle's stream
    gs = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    h.set_stream(gs)
    with torch.cuda.graph(graph, stream=gs):
        run()
    h.set_stream(torch.cuda.current_stream())
    times = []
    for k in range(a.warmup + a.steps):
        work.copy_(a0); bw.copy_(b0)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        with torch.cuda.stream(gs):
            e0.record(gs)
            graph.replay()
            e1.record(gs)
        torch.cuda.synchronize()
        if k >= a.warmup:
            times.append(e0.elapsed_time(e1))
    t = sum(times) / len(times)
    bad = int((info != 0).sum().item())
    if bad:
        fail(f"batched: {bad} zero pivots")
    # accuracy: backward error of a sample of the solves (host FP64)
    samp = rng.choice(a.count, 8, replace=False)
    berr = 0.0
    for i in samp:
        n, off, bo = int(sizes[i]), int(offs[i]), int(boffs[i])
        am = a0[off:off + n * n].view(n, n).t().double().cpu().numpy()
        x = bw[bo:bo + n].double().cpu().numpy()
        bb = b0[bo:bo + n].double().cpu().numpy()
        r = bb - am @ x
        berr = max(berr, float(np.abs(r).max() / (np.abs(am).sum(0).max() * np.abs(x).max()
                                                  + np.abs(bb).max())))
    # per-matrix launches: one driver call (RL, nb = ib = 32) + one solve each
    m = min(a.per_matrix_sample, a.count)
    pick = np.linspace(0, a.count - 1, m).astype(int)
    cfg = rf.api.make_config(W=rf.R_32F, I=rf.TC32, O=rf.TC32, ib=32)
    work.copy_(a0); bw.copy_(b0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in pick:
        n, off, bo = int(sizes[i]), int(offs[i]), int(boffs[i])
        view = work[off:off + n * n].view(n, n).t()
        rf.lu_nopiv(h, view, 32, algo=rf.RL, cfg=cfg)
        rf.getrs_nopiv(h, view, bw[bo:bo + n].view(1, n).t())
    torch.cuda.synchronize()
    t_single = (time.perf_counter() - t0) * 1e3
    scale = flops / float((2.0 * sizes[pick].astype(np.float64) ** 3 / 3.0).sum())
    gbs = bytes_ / (t * 1e-3) / 1e9
    tfl = flops / (t * 1e-3) / 1e12
    return {
        "metric": "batched no-pivot LU + solve: 20000 matrices, n ~ U[16,512] (GFLOP/s)",
        "value": round(flops / (t * 1e-3) / 1e9, 1), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 storage, fp16 MFMA updates (fp32 accumulate)",
        "data": "synthetic: seeded minstd_rand0 uniform(-1,1), a_ii = 1.2*sum|a_ij|",
        "config": {"workload": f"{a.count} LUs + solves: one persistent LU launch per size "
                               "bucket (autotuned shape) and one solve launch, in one HIP graph",
                   "count": a.count, "seed": a.seed, "panel": 32,
                   "buckets": [[bm, i1 - i0] for i0, i1, bm in buckets]},
        "flops": flops, "matrices_per_s": round(a.count / (t * 1e-3), 1),
        "roofline": {"bound": "hbm", "algorithmic_bytes": bytes_,
                     "work": "sum 8 n_i^2 bytes (fp32 read + write once)",
                     "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4),
                     "hbm_floor_ms": round(bytes_ / (PEAK_HBM_GBS * 1e9) * 1e3, 3),
                     "mfma_tflops": round(tfl, 2),
                     "mfma_frac": round(tfl / PEAK_FP16_TFLOPS, 5)},
        "zero_pivots": bad, "sample_backward_error": berr,
        "per_matrix_launches": {"sample": int(m), "ms_sample": round(t_single, 2),
                                "ms_extrapolated": round(t_single * scale, 1),
                                "speedup_of_batched": round(t_single * scale / t, 1)},
    }


# ---------------------------------------------------------------- config 5
def bench_autograd(a):
    """X = A^-1 B as a torch.autograd.Function: A 8192^2 fp32 (alpha = 1.1),
    B 8192 x 64, FP16-MFMA factorization + FP64-residual refinement, then the
    adjoint (transposed solves with the stored factors, gA = -gB X^T)."""
    import torch

    import remifa_amd as rf
    from remifa_amd import autograd as ag

    dev = torch.device("cuda:0")
    n, k = 8192, 64
    h = rf.Handle()
    Ac = rf.colmajor_empty(n, n, device=dev)
    rf.gen_diagdom(h, Ac, seed=a.seed, alpha=1.1)
    Bc = rf.colmajor_empty(n, k, device=dev)
    rf.gen_uniform(h, Bc, seed=a.seed + 1)
    A = Ac.contiguous().requires_grad_()
    B = Bc.double().contiguous().requires_grad_()
    W = torch.ones(n, k, dtype=torch.float64, device=dev)
    times = []
    for i in range(a.warmup + max(a.steps, 2)):
        A.grad = None
        B.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X = ag.solve(A, B, nb=256)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (X * W).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if i >= a.warmup:
            times.append(((t1 - t0) * 1e3, (t2 - t1) * 1e3))
    fw = sum(t[0] for t in times) / len(times)
    bw = sum(t[1] for t in times) / len(times)
    Xc = rf.colmajor_empty(n, k, dtype=torch.float64, device=dev)
    Xc.copy_(X.detach())
    Bcd = rf.colmajor_empty(n, k, dtype=torch.float64, device=dev)
    Bcd.copy_(B.detach())
    m = rf.error_measures(h, Ac, Xc, Bcd)
    # the adjoint: gB = A^-T W solves A^T gB = W; its backward error
    gB = rf.colmajor_empty(n, k, dtype=torch.float64, device=dev)
    gB.copy_(B.grad)
    Wc = rf.colmajor_empty(n, k, dtype=torch.float64, device=dev)
    Wc.copy_(W)
    At = rf.colmajor_empty(n, n, device=dev)
    At.copy_(Ac.t())
    mt = rf.error_measures(h, At, gB, Wc)
    flops = 2.0 * n**3 / 3.0
    return {"metric": "differentiable X = A^-1 B, A 8192^2, B 8192x64: forward + backward ms",
            "value": round(fw + bw, 3), "unit": "ms", "n_gpus": 1, "steps": len(times),
            "warmup": a.warmup, "ms_per_step": round(fw + bw, 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 storage, fp16 MFMA updates; fp64 residual",
            "data": "synthetic: seeded minstd_rand0, a_ii = 1.1*sum|a_ij|",
            "config": {"workload": "autograd solve n=8192 nrhs=64", "nb": 256},
            "forward_ms": round(fw, 3), "backward_ms": round(bw, 3),
            "factor_tflops_upper_bound": round(flops / (fw * 1e-3) / 1e12, 2),
            "backward_error": m["nwerr"], "adjoint_backward_error": mt["nwerr"]}


# ---------------------------------------------------------------- config 2
def bench_c2(a):
    """Config 2: fp32 storage with TC32 outer updates (lu_nopiv_ll_f32_fp32xtc32,
    src/lu_nopiv_ll.cxx:74-87, ib = 8), n = 16384, nb = 512, alpha in {1.01,
    1.1, 2}; beside it the plain FP32 factorization (lu_nopiv_ll_f32) of the
    same matrix. For each: factor time, the reference harness's check solve
    (b = A 1, fp32 solve with identity pivots) with its errors against FP64
    (normwise, componentwise, forward error against x = 1), the factor
    error, and the FP64-residual refinement to 1e-14."""
    import torch

    import remifa_amd as rf

    dev = torch.device("cuda:0")
    n, nb = 16384, 512
    h = rf.Handle()
    A = rf.colmajor_empty(n, n, device=dev)
    W = rf.colmajor_empty(n, n, device=dev)
    cpy = rf.colmajor_empty(n, n, dtype=torch.float16, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    flops = 2.0 * n**3 / 3.0
    out = {"workload": "fp32 storage LU, TC32 outer updates (lu_nopiv_ll_f32_fp32xtc32) vs "
                       "plain FP32 (lu_nopiv_ll_f32)", "n": n, "nb": nb, "ib": 8,
           "seed": a.seed, "rhs": "b = A 1 (FP64), x* = 1", "alphas": {}}
    for alpha in (1.01, 1.1, 2.0):
        rf.gen_diagdom(h, A, seed=a.seed, alpha=alpha)
        B = (A.double() @ torch.ones(n, 1, dtype=torch.float64, device=dev))
        Bc = rf.colmajor_empty(n, 1, dtype=torch.float64, device=dev)
        Bc.copy_(B)
        ones = torch.ones(n, 1, dtype=torch.float64, device=dev)
        res = {}
        for name, fn, with_cpy in (("fp32xtc32", rf.lu_nopiv_ll_f32_fp32xtc32, True),
                                   ("fp32", rf.lu_nopiv_ll_f32, False)):
            ts = []
            for k in range(2):
                W.copy_(A)
                torch.cuda.synchronize()
                e0, e1 = ev(), ev()
                e0.record()
                if with_cpy:
                    fn(h, n, n, nb, W, n, info, cpy)
                else:
                    fn(h, n, n, nb, W, n, info)
                e1.record()
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            inf = int(info.item())
            r = {"factor_ms": round(ts[-1], 3),
                 "tflops": round(flops / (ts[-1] * 1e-3) / 1e12, 2), "info": inf}
            if inf != 0:
                fail(f"c2 {name} alpha={alpha}: info {inf}")
                res[name] = r
                continue
            xs = rf.colmajor_empty(n, 1, device=dev)
            xs.copy_(Bc)
            rf.getrs_nopiv(h, W, xs)
            X = xs.double()
            r["solve"] = measures(h, A, X, Bc, W, xref=ones)
            X2 = torch.zeros((1, n), dtype=torch.float64, device=dev).t()
            it, be, _ = rf.ir_solve(h, A, W, Bc, X2, tol=1e-14, maxit=50)
            r["ir_iters_to_1e-14"] = it
            r["ir_backward_error"] = be
            r["ir_forward_error"] = float((X2 - ones).abs().max())
            if it < 0:
                fail(f"c2 {name} alpha={alpha}: refinement did not converge")
            res[name] = r
        if "solve" in res.get("fp32xtc32", {}) and "solve" in res.get("fp32", {}):
            res["nwerr_ratio_tc32_over_fp32"] = (res["fp32xtc32"]["solve"]["nwerr"]
                                                 / res["fp32"]["solve"]["nwerr"])
            res["factor_speedup"] = round(res["fp32"]["factor_ms"]
                                          / res["fp32xtc32"]["factor_ms"], 3)
        out["alphas"][str(alpha)] = res
    del A, W, cpy
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------- headline
def bench_headline(a):
    import numpy as np
    import torch

    import remifa_amd as rf

    dev = torch.device("cuda:0")
    if a.pmc_calibrate:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst
    n, nb = a.n, a.nb
    h = rf.Handle()
    h.set_lookahead(not a.no_lookahead, a.reserve_cus)
    s = h.stream
    A = rf.colmajor_empty(n, n, device=dev)
    rf.gen_diagdom(h, A, seed=a.seed, alpha=a.alpha)
    bf = rf.colmajor_empty(n, a.nrhs, device=dev)
    rf.gen_uniform(h, bf, seed=a.seed + 1)
    B = rf.colmajor_empty(n, a.nrhs, dtype=torch.float64, device=dev)
    B.copy_(bf)
    X = rf.colmajor_empty(n, a.nrhs, dtype=torch.float64, device=dev)
    A16 = rf.colmajor_empty(n, n, dtype=torch.float16, device=dev)
    algo = rf.LL if a.algo == "ll" else rf.RL
    o_upd = {"fp16": rf.TC32, "bf16": rf.BF16_TC32, "fp8": rf.FP8_TC32}[a.update]
    if algo == rf.LL:
        cfg16 = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=o_upd, OP=rf.R_32F, ib=a.ib)
    else:
        cfg16 = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=o_upd, ib=a.ib)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def step16(timing: bool, cfg=cfg16):
        rf.convert(h, A, A16)  # restore the working copy (untimed)
        X.zero_()
        info.fill_(-1)
        e0, e1, e2 = ev(), ev(), ev()
        if timing:
            h.timing(True)
        e0.record(s)
        rf.lu_nopiv(h, A16, nb, algo=algo, cfg=cfg, info=info)
        e1.record(s)
        it, be, hist = rf.ir_solve(h, A, A16, B, X, tol=1e-14, maxit=50)
        e2.record(s)
        h.timing(False)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2), it, be, int(info.item())

    def checked(r, what):
        if r[4] != 0:
            fail(f"{what}: info {r[4]}")
        if r[2] < 0:
            fail(f"{what}: refinement did not reach 1e-14 (backward error {r[3]:.3e})")
        return r

    for _ in range(a.warmup):
        checked(step16(False), "warmup step")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    res = [step16(False) for _ in range(a.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    for r in res:
        checked(r, "timed step")
    t_fac = sum(r[0] for r in res) / len(res)
    t_ir = sum(r[1] for r in res) / len(res)
    iters, nwerr = res[-1][2], res[-1][3]
    # per-kernel HIP events (two per launch, ~2300 launches per step) would
    # add ~15% to the timed steps, so they bracket the launches of a second,
    # instrumented pass of the same steps: the breakdown and the roofline's
    # per-launch durations come from it
    h.timing_reset()
    res_i = [checked(step16(True), "instrumented step") for _ in range(a.steps)]
    t_fac_inst = sum(r[0] for r in res_i) / len(res_i)
    gemm_ms, gemm_n, gemm_flops = h.timing_query(0)
    panel_ms, panel_n, panel_bytes = h.timing_query(1)
    conv_ms, conv_n, conv_bytes = h.timing_query(2)
    trsm_ms, trsm_n, _ = h.timing_query(3)
    res_ms, res_n, _ = h.timing_query(4)
    # dominant kernel instantiation: the GEMM variant with the most time
    cls, var, kms, kwork = h.timing_records()
    g = cls == 0
    variants = {}
    for v in np.unique(var[g]):
        sel = g & (var == v)
        variants[int(v)] = (kms[sel].sum(), int(sel.sum()), kwork[sel].sum())
    dom = max(variants, key=lambda v: variants[v][0])
    dom_ms, dom_n, dom_flops = variants[dom]
    h.timing_reset()

    # HBM traffic of the dominant kernel from the committed --pmc summary of
    # this workload (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated)
    traffic, traffic_src = None, None
    bm, bn = rf.api._GEMM_TILE.get(dom & 15, (0, 0, 0, 0))[:2]
    kname = "gemm_big" if (dom & 15) == 4 else "gemm_mfma"
    prof = os.path.join(ROOT, "profiles", f"{kname}_{bm}x{bn}_lu_n{n}_pmc.md")
    if os.path.exists(prof):
        import re
        m = re.search(r"traffic_bytes_per_launch = (\d+)", open(prof).read())
        if m:
            traffic, traffic_src = int(m.group(1)), os.path.relpath(prof, ROOT)

    # algorithmic bytes of the dominant kernel's launches among the wide
    # updates of LL: A and B in fp16 once, the fp32 panel buffer C read and
    # written once. The 256x256 kernel takes the launches with K >= 2048, the
    # 128x128 one the shorter ones (every one for the FP8 kernel).
    algo_bytes = None
    if algo == rf.LL:
        tot, cnt = 0, 0
        # (with the lookahead block k's update is two launches of each shape:
        # K = ofs - nb on the second stream, then K = nb for block k-1)
        la = not a.no_lookahead and (a.update != "fp8" or nb % 128 == 0)
        dom_big, dom_fp8 = (dom & 15) == 4, (dom >> 8) == 3
        for k in range(1, (n + nb - 1) // nb):
            ofs = k * nb
            ks = ([ofs - nb] if k >= 2 else []) + [nb] if la else [ofs]
            for K in ks:
                if not dom_fp8 and (K >= 2048) != dom_big:
                    continue
                for (M, N) in ((n - ofs, min(nb, n - ofs)), (min(nb, n - ofs), n - ofs)):
                    tot += 2 * (M * K + K * N) + 8 * M * N
                    cnt += 1
        algo_bytes = round(tot / cnt) if cnt else None

    flops = 2.0 * n**3 / 3.0
    gemm_tflops = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    dom_tflops = dom_flops / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
    upd_name = {"fp16": "TC32", "bf16": "BF16_TC32", "fp8": "FP8_TC32"}[a.update]
    out = {
        "metric": METRIC,
        "value": round(flops / (t_fac * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_fac + t_ir, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp16": "fp16", "bf16": "bf16 outer / fp16 inner",
                  "fp8": "e4m3 (per-32-block scaled) outer / fp16 inner"}[a.update]
                 + " (fp32 accumulate); fp64 residual",
        "data": f"synthetic: seeded minstd_rand0 uniform(-1,1), a_ii = {a.alpha}*sum|a_ij|, "
                f"generated on the GPU",
        "config": {"workload": f"no-pivot LU n={n} fp16 storage, {upd_name} MFMA outer updates, "
                               f"+ FP64-residual IR to 1e-14, nrhs={a.nrhs}",
                   "n": n, "nb": nb, "ib": a.ib, "algo": a.algo, "alpha": a.alpha,
                   "nrhs": a.nrhs, "seed": a.seed,
                   "lookahead": algo == rf.LL and not a.no_lookahead
                   and (a.update != "fp8" or nb % 128 == 0),
                   "lookahead_reserved_cus": a.reserve_cus,
                   "precision": f"W=f16 I=TC32 O={upd_name}"
                                + (" OP=IP=OF=IF=f32" if algo == rf.LL else "")},
        "factor_ms": round(t_fac, 3),
        "ir_ms": round(t_ir, 3),
        "ir_iters": iters,
        "backward_error": nwerr,
        "wall_s_timed_steps": round(wall, 3),
        "factor_ms_instrumented_pass": round(t_fac_inst, 3),
        "kernels": {
            "gemm": {"ms": round(gemm_ms / a.steps, 3), "launches": gemm_n // a.steps,
                     "tflops": round(gemm_tflops, 2)},
            "panel": {"ms": round(panel_ms / a.steps, 3), "launches": panel_n // a.steps,
                      "gbps": round(panel_bytes / (panel_ms * 1e-3) / 1e9, 1) if panel_ms else 0},
            "convert": {"ms": round(conv_ms / a.steps, 3), "launches": conv_n // a.steps,
                        "gbps": round(conv_bytes / (conv_ms * 1e-3) / 1e9, 1) if conv_ms else 0},
            "trsm": {"ms": round(trsm_ms / a.steps, 3), "launches": trsm_n // a.steps},
            "residual": {"ms": round(res_ms / a.steps, 3), "launches": res_n // a.steps},
        },
        "roofline": {"bound": "mfma", "achieved": round(dom_tflops, 2),
                     "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(dom_tflops / PEAK_FP16_TFLOPS, 4), "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM read + write, rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "kernel": rf.api.gemm_variant_name(dom),
                     "launches_per_step": dom_n // a.steps,
                     "avg_launch_us": round(dom_ms / dom_n * 1e3, 2),
                     "flops_per_launch": dom_flops / dom_n,
                     "work_per_launch": "2*M*N*K flops of the launch's update GEMM",
                     "all_update_gemms_tflops": round(gemm_tflops, 2)},
        "gemm_variants": {rf.api.gemm_variant_name(v): {
            "ms_per_step": round(t[0] / a.steps, 3), "launches_per_step": t[1] // a.steps,
            "tflops": round(t[2] / (t[0] * 1e-3) / 1e12, 2) if t[0] else 0}
            for v, t in variants.items()},
    }

    # errors of the headline result: the FP64 reference solution (rocSOLVER
    # FP64 no-pivot LU + solve), the GPU measures, the host recheck
    xref = None
    if not a.no_errors:
        try:
            xref = fp64_reference_solution(A, B)
        except Exception as e:  # the comparator library is optional
            out["fp64_reference_solution"] = f"unavailable: {e}"
        out["errors"] = measures(h, A, X, B, A16, xref=xref)
        x0 = rf.colmajor_empty(n, a.nrhs, device=dev)
        x0.copy_(bf)
        rf.getrs_nopiv(h, A16, x0)  # the unrefined fp32 solve from the fp16 factors
        out["errors"]["unrefined_solve"] = measures(h, A, x0.double(), B, A16, factor=False,
                                                    xref=xref)
        hb = host_nwerr(A, B[:, 0].cpu().numpy(), X[:, 0].cpu().numpy())
        out["errors"]["host_fp64_nwerr"] = hb
        if not hb <= 1e-14:
            fail(f"host FP64 recheck of the headline solution: nwerr {hb:.3e} > 1e-14")

    t32 = None
    if not a.no_fp32:
        # plain FP32 GPU factorization (W=f32, I=O=FP32), same matrix, LL and
        # RL; the faster is the comparator
        A32 = rf.colmajor_empty(n, n, device=dev)
        cfg32 = rf.api.make_config(W=rf.R_32F, I=rf.FP32, O=rf.FP32, ib=a.ib)
        nb32 = a.fp32_nb or nb
        bsol = rf.colmajor_empty(n, a.nrhs, device=dev)
        runs = {}
        for alg_name, alg in (("ll", rf.LL), ("rl", rf.RL)):
            t32 = []
            for k in range(a.fp32_steps + 1):
                A32.copy_(A)
                torch.cuda.synchronize()
                e0, e1, e2, e3 = ev(), ev(), ev(), ev()
                e0.record(s)
                inf = rf.lu_nopiv(h, A32, nb32, algo=alg, cfg=cfg32)
                e1.record(s)
                bsol.copy_(bf)
                rf.getrs_nopiv(h, A32, bsol)
                e2.record(s)
                X32 = torch.zeros_like(X)
                it32, be32, _ = rf.ir_solve(h, A, A32, B, X32, tol=1e-14, maxit=50)
                e3.record(s)
                torch.cuda.synchronize()
                if inf != 0:
                    fail(f"fp32 {alg_name}: info {inf}")
                if it32 < 0:
                    fail(f"fp32 {alg_name}: refinement did not converge")
                if k > 0 or a.fp32_steps == 0:
                    t32.append((e0.elapsed_time(e1), e1.elapsed_time(e2), e2.elapsed_time(e3),
                                it32))
            lu32 = sum(t[0] for t in t32) / len(t32)
            r = {"factor_ms": round(lu32, 3),
                 "tflops": round(flops / (lu32 * 1e-3) / 1e12, 3),
                 "solve_ms": round(sum(t[1] for t in t32) / len(t32), 3),
                 "ir_ms": round(sum(t[2] for t in t32) / len(t32), 3),
                 "ir_iters": t32[-1][3], "nb": nb32}
            if not a.no_errors:
                r["errors"] = measures(h, A, bsol.double(), B, A32, xref=xref)
            runs[alg_name] = r
        best = min(runs, key=lambda k: runs[k]["factor_ms"])
        lu32, sv32, ir32 = (runs[best]["factor_ms"], runs[best]["solve_ms"], runs[best]["ir_ms"])
        tts = t_fac + t_ir
        out["fp32"] = dict(runs[best], algo=best, other={k: v for k, v in runs.items()
                                                         if k != best})
        out["fp32"]["roofline_frac_of_fp32_peak"] = round(
            flops / (lu32 * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4)
        out["speedup_tts_vs_fp32_lu_fp32_solve"] = round((lu32 + sv32) / tts, 3)
        out["speedup_tts_vs_fp32_lu_fp64_ir"] = round((lu32 + ir32) / tts, 3)
        out["speedup_factor_vs_fp32"] = round(lu32 / t_fac, 3)
        t32 = lu32
        del A32

    if not a.no_vendor:
        try:
            out["vendor_rocsolver"] = bench_vendor(h, A, B, xref, t32)
        except Exception as e:
            out["vendor_rocsolver"] = f"unavailable: {e}"

    if not a.no_variants:
        # the same factorization with BF16 and FP8 (per-32-block scaled e4m3)
        # outer updates, then FP64-residual refinement to 1e-14; and the
        # nrhs = 16 refinement on the FP16 factors
        var = {}
        for name, O in (("bf16", rf.BF16_TC32), ("fp8", rf.FP8_TC32)):
            if algo == rf.LL:
                cfgv = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=O, OP=rf.R_32F, ib=a.ib)
            else:
                cfgv = rf.api.make_config(W=rf.R_16F, I=rf.TC32, O=O, ib=a.ib)
            res_v = [checked(step16(False, cfgv), f"{name} variant") for _ in range(2)]
            fv, iv, itv, bev, _ = res_v[-1]
            var[name] = {"factor_ms": round(fv, 3), "tflops": round(flops / (fv * 1e-3) / 1e12, 2),
                         "ir_ms": round(iv, 3), "ir_iters": itv, "backward_error": bev,
                         "time_to_solution_ms": round(fv + iv, 3)}
            if not a.no_errors:
                var[name]["errors"] = measures(h, A, X, B, A16, xref=xref)
        checked(step16(False), "nrhs=16 setup")  # fp16 factors back in A16
        b16f = rf.colmajor_empty(n, 16, device=dev)
        rf.gen_uniform(h, b16f, seed=a.seed + 2)
        B16 = rf.colmajor_empty(n, 16, dtype=torch.float64, device=dev)
        B16.copy_(b16f)
        X16 = torch.zeros_like(B16)
        rf.ir_solve(h, A, A16, B16, X16, tol=1e-14, maxit=30)  # warmup
        X16.zero_()
        e0, e1 = ev(), ev()
        e0.record(s)
        it16, be16, _ = rf.ir_solve(h, A, A16, B16, X16, tol=1e-14, maxit=30)
        e1.record(s)
        torch.cuda.synchronize()
        if it16 < 0:
            fail("nrhs=16 refinement did not converge")
        var["fp16_nrhs16"] = {"ir_ms": round(e0.elapsed_time(e1), 3), "ir_iters": it16,
                              "backward_error": be16}
        if not a.no_errors:
            var["fp16_nrhs16"]["errors"] = measures(h, A, X16, B16, A16, factor=False)
        out["update_variants"] = var
        del b16f, B16, X16
    del A, A16, X, B
    torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); sharding the "
              "factorization over several GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batched":
        print(json.dumps(bench_batched(a)), flush=True)
    elif a.workload == "autograd":
        print(json.dumps(bench_autograd(a)), flush=True)
    elif a.workload == "c2":
        print(json.dumps(bench_c2(a)), flush=True)
    else:
        out = bench_headline(a)
        if not a.no_configs:
            out["configs"] = {"c2": bench_c2(a), "c4_batched": bench_batched(a),
                              "c5_autograd": bench_autograd(a)}
        if not a.no_cpu:
            out["cpu_baseline"] = cpu_baseline(a.cpu_seconds, a.cpu_n, a.seed, a.alpha)
        if FAILURES:
            out["failures"] = FAILURES
        print(json.dumps(out), flush=True)
    sys.exit(3 if FAILURES else 0)


if __name__ == "__main__":
    main()
