"""Writes the golden fixtures under tests/golden/.

  minstd.json   -- values of libstdc++'s std::default_random_engine with
                   std::uniform_real_distribution<float>(-1, 1), the generator
                   of the reference's gen_diagdom (tests/genmat.hxx:72-73),
                   drawn by minstd_ref.cpp compiled with the host g++;
  convert.json  -- narrowing truth table (fp32 -> fp16 / bf16 / e4m3 bits) on
                   boundary values, from numpy's float16 and torch's bfloat16 /
                   float8_e4m3fn casts (independent of the oracle);
  ir_trace.json -- backward error per refinement step of the oracle for one
                   seeded n=256 case (a regression fixture of the oracle);
  generators.json -- libstdc++ uniform_real_distribution<float>(0, 1) draws
                   and FNV-1a hashes of the reference's gen_arrowhead and
                   gen_randmat matrices (tests/common.hxx:298-319, 391-455)
                   built by minstd_ref.cpp with the real std:: engine;
  geequb.json   -- LAPACK sgeequb (scipy's bundled LAPACK) on seeded
                   badly scaled matrices: r, c, rowcnd, colcnd, amax, info.

usage: python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

MINSTD_CASES = [(-1, 512), (1, 64), (7, 64), (2024, 64), (123456789, 64)]


def minstd_fixture():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "minstd_ref")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "minstd_ref.cpp"),
                        "-o", exe], check=True)
        out = {}
        for seed, count in MINSTD_CASES:
            r = subprocess.run([exe, str(seed), str(count)], check=True, capture_output=True,
                               text=True)
            out[str(seed)] = r.stdout.split()
    return out


ARROW_CASES = [(300, 0.0, 1.0), (300, -1.0, 1.0), (600, 0.0, 1.0), (200, -1.0, 1.0)]
RANDMAT_CASES = [(37, 29, 5), (256, 256, 1), (1000, 300, 1)]


def generators_fixture():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "minstd_ref")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "minstd_ref.cpp"),
                        "-o", exe], check=True)
        run = lambda *a: subprocess.run([exe, *map(str, a)], check=True, capture_output=True,
                                        text=True).stdout.split()
        return {
            "uniform_0_1": run("uniform", 0.0, 1.0, 64),
            "arrowhead": [[n, lo, hi, run("arrowhead", n, lo, hi)[0]] for n, lo, hi in ARROW_CASES],
            "randmat": [[m, n, s, run("randmat", m, n, s)[0]] for m, n, s in RANDMAT_CASES],
        }


def geequb_fixture():
    import scipy.linalg.lapack as lapack
    rng = np.random.default_rng(9)
    cases = []
    for m, n in ((40, 40), (17, 33), (64, 50)):
        a = rng.uniform(-1, 1, (m, n)) * 10.0 ** rng.integers(-6, 7, (m, 1))
        a *= 10.0 ** rng.integers(-3, 4, (1, n))
        a = np.asfortranarray(a.astype(np.float32))
        r, c, rowcnd, colcnd, amax, info = lapack.sgeequb(a)
        cases.append({"a_bits": [int(v) for v in a.ravel(order="F").view(np.uint32)],
                      "m": m, "n": n, "r_bits": [int(v) for v in r.view(np.uint32)],
                      "c_bits": [int(v) for v in c.view(np.uint32)], "rowcnd": float(rowcnd),
                      "colcnd": float(colcnd), "amax": float(amax), "info": int(info)})
    return {"source": "scipy.linalg.lapack.sgeequb", "cases": cases}


def boundary_values():
    f16_max = 65504.0
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1.5, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25,
            3 * 2.0 ** -26, 2.0 ** -26, 2.0 ** -25 * 1.0000001, f16_max, 65519.0, 65520.0,
            -65520.0, 1e10, -1e10, float("inf"), float("-inf"), 1 + 2.0 ** -11,
            1 + 3 * 2.0 ** -11, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 448.0, 464.0, 480.0, 500.0,
            2.0 ** -6, 2.0 ** -9, 2.0 ** -10, 3 * 2.0 ** -11, 0.1, -0.3, 3.14159265, 1e-30,
            1.17549435e-38, 1.4e-45]
    rng = np.random.default_rng(5)
    vals += list(rng.standard_normal(64) * 10.0 ** rng.integers(-8, 6, 64))
    return np.array(vals, np.float32)


def convert_fixture():
    import torch
    x = boundary_values()
    f16 = x.astype(np.float16).view(np.uint16)
    bf16 = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    # e4m3: in range only (the oracle saturates, torch's cast gives NaN beyond 448)
    sel = np.isfinite(x) & (np.abs(x) <= 464.0)
    e4 = torch.from_numpy(x[sel].copy()).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    return {
        "f32_bits": [int(v) for v in x.view(np.uint32)],
        "f16_bits": [int(v) for v in f16],
        "bf16_bits": [int(v) for v in bf16],
        "e4m3_f32_bits": [int(v) for v in x[sel].view(np.uint32)],
        "e4m3_bits": [int(v) for v in e4],
    }


def ir_fixture():
    sys.path.insert(0, ROOT)
    import oracle as orc
    n, seed, alpha = 256, 11, 1.05
    a = orc.gen_diagdom_alpha(n, seed, alpha)
    a16 = orc.convert(a, orc.F32, orc.F16)
    c = orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32, OP=orc.F32, ib=8)
    lu = orc.convert(orc.lu_nopiv("ll", c, a16, 64), orc.F16, orc.F32)
    b = orc.gen_uniform(n, 1, seed + 1).astype(np.float64)
    x, it, hist = orc.ir_solve(a, lu, b, tol=1e-14, maxit=30)
    return {"n": n, "seed": seed, "alpha": alpha, "nb": 64, "ib": 8,
            "config": "W=f16 I=TC32 O=TC32 OP=f32 (LL)", "iters": it,
            "history": [float(h) for h in hist]}


def main():
    with open(os.path.join(HERE, "minstd.json"), "w") as f:
        json.dump(minstd_fixture(), f, indent=0)
    with open(os.path.join(HERE, "convert.json"), "w") as f:
        json.dump(convert_fixture(), f)
    with open(os.path.join(HERE, "ir_trace.json"), "w") as f:
        json.dump(ir_fixture(), f, indent=1)
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(generators_fixture(), f, indent=1)
    with open(os.path.join(HERE, "geequb.json"), "w") as f:
        json.dump(geequb_fixture(), f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
