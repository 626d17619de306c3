#!/usr/bin/env python3
"""One runtime-dispatched LU test driver over the library's C ABI: the
reference's test binary (tests/testing_lu_gpu.cxx, lu_test in
tests/testing_lu_gpu.hxx:101-1431) with its options (tests/TestOpts.hxx:31-396).

    python tests/testing_lu_gpu.py --algo=remifa-ll --fp16 --in-upd=tc32 \\
        --out-upd=tc32 --out-pnl=fp32 --in-pnl=fp32 --m 8192 --nb 512 --diagdom

Flow, as lu_test: generate the test matrix in fp32 (on the GPU; --mtx reads
a MatrixMarket file on the host), optionally round it to fp16 (--cast-f16)
and equilibrate it (--scale: geequb + laqge), right-hand side b = A 1, then
factor with the selected algorithm and precision tuple, timed around the
driver call only (tests/testing_lu_gpu.hxx:514-531, 1017-1020), then check:
fp32 solve with identity pivots and the reference's error measures
(tests/errors.hxx) computed in FP64 on the GPU; --check-conv prints the
fp32 -> fp16 conversion error of A (tests/common.hxx:107-174). The printed
lines are the reference's ("nwerr = ...", "cwerr = ...", "factor time (s) =
...", "GFlop/s = ..."); --json prints one JSON object as well.

--algo=cusolver runs the vendor no-pivot LU (rocSOLVER sgetrf_npvt, the
counterpart of cuSOLVER getrf with a NULL pivot array, tests/testing_lu_gpu.hxx:
390-468). Generators: --diagdom (gen_diagdom), --rand (gen_randmat), --cond C
(gen_posdef_cond: Q D Q^T, D = diag(10^(-C i/(n-1))), Q from the QR of
gen_randmat, with rocSOLVER sgeqrf / sormqr like the reference's cuSOLVER
calls), --arrowhead / --arrowhead-pos / --arrowhead-pos-neg (gen_arrowhead),
--mtx FILE (read_mtx, tests/genmat.hxx:16-63), and the benchmark generator
--alpha A --seed S (a_ii = A sum_j!=i |a_ij|). Options of the reference that
need what this port does not build (LATMR / LATMS from LAPACK's testing
library, --fp64 working precision, CUTLASS, several GPUs) exit with a message.
Test harness only: it imports the library the way a user would; the oracle
is not used.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CTYPES = {"fp32": 0, "fp16": 1, "tc32": 2, "tc16": 3, "bf16tc32": 4, "fp8tc32": 5}
PRECS = {"fp32": 0, "fp16": 1}


def parser():
    p = argparse.ArgumentParser(prog="testing_lu_gpu", description=__doc__.split("\n\n")[0])
    p.add_argument("--algo", default="remifa",
                   choices=["remifa", "remifa-ll", "remifa-rl", "cusolver", "cusolver-hp"])
    p.add_argument("--m", "-m", type=int, default=256)
    p.add_argument("--n", "-n", type=int, default=256)
    p.add_argument("--k", "-k", type=int, default=256)
    p.add_argument("--nb", type=int, default=256)
    p.add_argument("--ib", type=int, default=8, help="inner block size (the reference fixes 8)")
    mat = p.add_mutually_exclusive_group()
    mat.add_argument("--rand", dest="mat", action="store_const", const="rand")
    mat.add_argument("--symrand", dest="mat", action="store_const", const="symrand")
    mat.add_argument("--posrand", dest="mat", action="store_const", const="posrand")
    mat.add_argument("--diagdom", dest="mat", action="store_const", const="diagdom")
    mat.add_argument("--diagdom-pos", dest="mat", action="store_const", const="diagdom-pos")
    mat.add_argument("--tridiagdom", dest="mat", action="store_const", const="tridiagdom")
    mat.add_argument("--arrowhead", dest="mat", action="store_const", const="arrowhead")
    mat.add_argument("--arrowhead-pos", dest="mat", action="store_const", const="arrowhead-pos")
    mat.add_argument("--arrowhead-pos-01", dest="mat", action="store_const",
                     const="arrowhead-pos-01")
    mat.add_argument("--arrowhead-pos-neg", dest="mat", action="store_const",
                     const="arrowhead-pos-neg")
    mat.add_argument("--cond", type=float, default=None)
    mat.add_argument("--mtx", default=None)
    mat.add_argument("--latmr", type=int, default=None)
    mat.add_argument("--latms", type=int, default=None)
    mat.add_argument("--alpha", type=float, default=None,
                     help="benchmark generator: a_ii = alpha * sum_j!=i |a_ij| (with --seed)")
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--gamma", type=float, default=1.0)
    p.add_argument("--ncpu", "-ncpu", type=int, default=1)
    p.add_argument("--ngpu", "-ngpu", type=int, default=1)
    p.add_argument("--enable-tc", dest="usetc", action="store_true", default=True)
    p.add_argument("--disable-tc", dest="usetc", action="store_false")
    p.add_argument("--no-check", dest="check", action="store_false", default=True)
    p.add_argument("--check-conv", action="store_true")
    prec = p.add_mutually_exclusive_group()
    prec.add_argument("--fp16", dest="prec", action="store_const", const="fp16")
    prec.add_argument("--fp32", dest="prec", action="store_const", const="fp32")
    prec.add_argument("--fp64", dest="prec", action="store_const", const="fp64")
    for side in ("in", "out"):
        p.add_argument(f"--{side}-upd", choices=list(CTYPES), default="fp32")
        p.add_argument(f"--{side}-pnl", choices=list(PRECS), default=None)
        p.add_argument(f"--{side}-fac", choices=list(PRECS), default=None)
    p.add_argument("--cutlass", action="store_true")
    p.add_argument("--cast-f16", action="store_true")
    p.add_argument("--scale", action="store_true")
    p.add_argument("--splitk", type=int, default=0, help="split-K slices of the updates")
    p.add_argument("--slv", choices=["fp32", "fp16"], default="fp32")
    p.add_argument("--no-lookahead", action="store_true")
    p.add_argument("--json", action="store_true")
    return p


def read_mtx(path: str):
    """tests/genmat.hxx:16-63: MatrixMarket coordinate file, '%' header and
    comment lines skipped, then 'nrows ncols nnz' and nnz 'row col value'
    triplets (1-based), each stored at (row, col) and at (col, row); the
    order is ncols. Returns a Fortran-order float32 numpy matrix."""
    import numpy as np
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    toks = " ".join(lines).split()
    nrows, ncols, nnz = int(toks[0]), int(toks[1]), int(toks[2])
    n = ncols
    a = np.zeros((n, n), np.float32, order="F")
    for c in range(nnz):
        r, col, v = int(toks[3 + 3 * c]) - 1, int(toks[4 + 3 * c]) - 1, float(toks[5 + 3 * c])
        a[r, col] = np.float32(v)
        a[col, r] = np.float32(v)
    return a, nrows


def gen_posdef_cond(h, n: int, cond: float):
    """tests/common.hxx:526-727 (mode 3): A = Q D Q^T, D = diag(10^(-cond i /
    (n - 1))), Q the orthogonal factor of the QR of gen_randmat (seed 1),
    applied with sgeqrf + sormqr (left, then right transposed) on the GPU"""
    import torch
    import remifa_amd as rf
    from bench import RocSolver
    dev = torch.device("cuda:0")
    lam = rf.colmajor_empty(n, n, device=dev)
    rf.gen_randmat(h, lam, seed=1)
    d = torch.pow(torch.tensor(10.0, dtype=torch.float64),
                  -cond * torch.arange(n, dtype=torch.float64) / max(n - 1, 1)).float()
    a = rf.colmajor_empty(n, n, device=dev)
    a.zero_()
    a.diagonal().copy_(d.to(dev))
    rs = RocSolver()
    tau = torch.zeros(n, dtype=torch.float32, device=dev)
    rs.sgeqrf(lam, tau)
    rs.sormqr(left=True, trans=False, qr=lam, tau=tau, c=a)
    rs.sormqr(left=False, trans=True, qr=lam, tau=tau, c=a)
    torch.cuda.synchronize()
    rs.close()
    return a


def generate(h, o):
    """the test matrix (fp32, column-major, on the GPU) per the options"""
    import numpy as np
    import torch
    import remifa_amd as rf
    dev = torch.device("cuda:0")
    m = o.m
    if o.mtx:
        a_host, _ = read_mtx(o.mtx)
        m = a_host.shape[0]
        a = rf.colmajor_empty(m, m, device=dev)
        a.copy_(torch.from_numpy(np.ascontiguousarray(a_host.T)).t())
        return a
    a = rf.colmajor_empty(m, m, device=dev)
    mat = o.mat or ("cond" if o.cond is not None or o.alpha is None else "alpha")
    if o.alpha is not None:
        rf.gen_diagdom(h, a, seed=o.seed, alpha=o.alpha)
    elif mat == "diagdom":
        rf.gen_diagdom(h, a)
    elif mat == "rand":
        rf.gen_randmat(h, a, seed=1)
    elif mat in ("arrowhead", "arrowhead-pos"):
        rf.gen_arrowhead(h, a, 0.0, 1.0)
    elif mat == "arrowhead-pos-neg":
        rf.gen_arrowhead(h, a, -1.0, 1.0)
    elif mat == "cond":
        a = gen_posdef_cond(h, m, 1.0 if o.cond is None else o.cond)
    else:
        raise SystemExit(f"[lu_test] Matrix generator NOT available: {mat}")
    return a


def run(argv=None) -> dict:
    import numpy as np
    import torch
    import remifa_amd as rf
    o = parser().parse_args(argv)
    if o.latmr is not None or o.latms is not None:
        raise SystemExit("[lu_test] LATMR / LATMS need LAPACK's testing library (not built)")
    if o.prec == "fp64":
        raise SystemExit("[lu_test] FP64 working precision is not built")
    if o.cutlass or o.ngpu != 1:
        raise SystemExit("[lu_test] CUTLASS GEMMs and several GPUs are not built")
    if o.algo == "cusolver-hp":
        raise SystemExit("[lu_test] cusolver-hp: no half-precision vendor LU")
    prec = o.prec or "fp32"
    dev = torch.device("cuda:0")
    h = rf.Handle()
    h.set_lookahead(not o.no_lookahead, 0)
    if o.splitk:
        h.set_tuning(0, o.splitk)
    out = {"algo": o.algo, "prec": prec}
    a = generate(h, o)
    m = a.shape[0]
    if o.cast_f16:  # round the generated matrix to fp16 values
        t16 = rf.colmajor_empty(m, m, dtype=torch.float16, device=dev)
        rf.convert(h, a, t16)
        rf.convert(h, t16, a)
    if o.scale:
        r = torch.zeros(m, dtype=torch.float32, device=dev)
        c = torch.zeros(m, dtype=torch.float32, device=dev)
        eq = rf.geequb(h, a, r, c)
        print(f"[lu_test] geequ info = {eq['info']}")
        print(f"[lu_test] geequ amax = {eq['amax']}")
        equed = rf.laqge(h, a, r, c, eq["rowcnd"], eq["colcnd"], eq["amax"])
        print(f"[lu_test] laqge equed = {equed}")
        out["equed"] = equed
    # b = A x with x = 1 (tests/testing_lu_gpu.hxx:316-351), here in FP64
    ones = torch.ones(m, 1, dtype=torch.float64, device=dev)
    zero = torch.zeros(m, 1, dtype=torch.float64, device=dev)
    nb_ = rf.colmajor_empty(m, 1, dtype=torch.float64, device=dev)
    rf.residual_f64(h, a, ones, zero, nb_)  # -A 1
    b = (-nb_).contiguous()
    bcol = rf.colmajor_empty(m, 1, dtype=torch.float64, device=dev)
    bcol.copy_(b)
    if o.check_conv:
        a16 = rf.colmajor_empty(m, m, dtype=torch.float16, device=dev)
        rf.convert(h, a, a16)
        ce = rf.conv_error(h, a, a16)
        print(f"conv cwerr (abs) = {ce['abs']:e}")
        print(f"conv cwerr = {ce['rel']:e}")
        out["conv_cwerr"], out["conv_cwerr_abs"] = ce["rel"], ce["abs"]
        del a16
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if o.algo == "cusolver":
        from bench import RocSolver
        if prec != "fp32":
            raise SystemExit("[lu_test] the vendor LU runs in fp32")
        f = rf.colmajor_empty(m, m, device=dev)
        f.copy_(a)
        rs = RocSolver()
        torch.cuda.synchronize()
        e0.record()
        rs.getrf_npvt(f, info)
        e1.record()
        torch.cuda.synchronize()
        rs.close()
    else:
        algo = rf.RL if o.algo == "remifa-rl" else rf.LL
        W = PRECS[prec]
        OP = PRECS[o.out_pnl] if o.out_pnl else W
        IP = PRECS[o.in_pnl] if o.in_pnl else OP
        OF = PRECS[o.out_fac] if o.out_fac else OP
        IF = PRECS[o.in_fac] if o.in_fac else IP
        cfg = rf.LuConfig(W, CTYPES[o.in_upd], CTYPES[o.out_upd], OP, IP, OF, IF, o.ib)
        if W == rf.R_16F:
            f = rf.colmajor_empty(m, m, dtype=torch.float16, device=dev)
            rf.convert(h, a, f)  # before the timer, as the reference
        else:
            f = rf.colmajor_empty(m, m, device=dev)
            f.copy_(a)
        need_cpy = W == rf.R_32F and (cfg.I != rf.FP32 or cfg.O != rf.FP32)
        cpy = rf.colmajor_empty(m, m, dtype=torch.float16, device=dev) if need_cpy else None
        torch.cuda.synchronize()
        e0.record()
        rf.lu_nopiv(h, f, o.nb, algo=algo, cfg=cfg, info=info, a_cpy=cpy)
        e1.record()
        torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3
    out["info"] = int(info.item())
    if out["info"] != 0:
        print(f"[lu_test][error] zero or non-finite pivot at column {out['info']}")
    if o.check and out["info"] == 0:
        x = rf.colmajor_empty(m, 1, device=dev)
        x.copy_(bcol)
        rf.getrs_nopiv(h, f, x)
        xd = x.double()
        em = rf.error_measures(h, a, xd, bcol, lu=f)
        print(f"nwerr = {em['nwerr']:e}")
        print(f"cwerr = {em['cwerr']:e}")
        out.update(nwerr=em["nwerr"], cwerr=em["cwerr"], cwerr_oettli_prager=em["cwerr_ax"],
                   fwderr=float((xd - 1.0).abs().max()))
    flops = 2.0 * m**3 / 3.0
    print(f"factor time (s) = {t:e}")
    print(f"GFlop/s = {flops / (t * 1e9):.3f}")
    out.update(m=m, nb=o.nb, factor_s=t, gflops=flops / (t * 1e9))
    if o.json:
        print(json.dumps(out))
    return out


if __name__ == "__main__":
    run()
