"""Autotuner for the LU drivers' update GEMMs and slice width.

The drivers issue three families of matrix-core GEMMs (C -= A B, fp16 or
bf16 operands, fp32 accumulation):
  * narrow N (N <= 64): the inner slices' column updates and the RL rank-64
    updates;
  * narrow M (M <= 64): their row counterparts;
  * wide: the outer block-column / block-row updates (LL, N or M = nb) and
    the RL Schur complement.
For each family the tuner times every candidate (tile shape x split-K) on
the shapes that family takes in a factorization of order n with block nb,
keeps the fastest, and writes a rule table (JSON). A Handle loads the table
at creation (remifa_set_gemm_rules) unless REMIFA_TUNING=off; the table
committed next to this file was measured on an MI355X.

It also picks, per size bucket of the batched factorization, the thread
count and persistent workgroups per CU (remifa_set_batched_tuning).

usage (on a GPU): python -m remifa_amd.autotune [--n 32768] [--nb 512] [--out FILE]
                  [--batched-only]
"""
from __future__ import annotations

import argparse
import json
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_gfx950.json")

# tile ids (remifa_set_gemm_tile): 0 256x32, 1 32x256, 2 256x64, 3 64x256,
# 4 256x256 (LDS-DMA kernel), 5 128x128, 6 256x128, 7 128x256
FAMILIES = {
    "narrow_n": {"tiles": [0, 2, 5], "splits": [0, 1, 2, 4, 8]},
    "narrow_m": {"tiles": [1, 3, 5], "splits": [0, 1, 2, 4, 8]},
    "wide": {"tiles": [4, 5, 6, 7], "splits": [0, 1, 2]},
}


def driver_shapes(n: int, nb: int, pw: int = 64, samples: int = 16):
    """(family, M, N, K) of representative GEMMs of the LL driver at order n:
    the updates of `samples` outer blocks spread evenly over the
    factorization (the best split-K and tile change along it: early updates
    have many output tiles and short K, late ones few tiles and long K)"""
    out = []
    nblk = (n + nb - 1) // nb
    step = max(1, (nblk - 1) // samples)
    for kblk in range(1, nblk, step):
        ofs = kblk * nb
        m = n - ofs
        out.append(("wide", m, nb, ofs))          # block column (:239-304)
        out.append(("wide", nb, m - nb, ofs))     # block row (:311-410)
        for so in (pw, nb - pw):
            cm = m - so
            out.append(("narrow_n", cm, pw, so))  # slice columns
            out.append(("narrow_m", pw, cm - pw, so))  # slice rows
    return out


def _time_gemm(h, rf, torch, M, N, K, tile, split, reps=3):
    dev = torch.device("cuda:0")
    a = rf.colmajor_empty(M, K, dtype=torch.float16, device=dev)
    b = rf.colmajor_empty(K, N, dtype=torch.float16, device=dev)
    c = rf.colmajor_empty(M, N, device=dev)
    a.uniform_(-1, 1)
    b.uniform_(-1, 1)
    c.zero_()
    h.set_gemm_tile(tile)
    h.set_tuning(0, split)
    try:
        rf.gemmex(h, rf.TC32, a, b, c)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(h.stream)
        for _ in range(reps):
            rf.gemmex(h, rf.TC32, a, b, c)
        e1.record(h.stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps
    finally:
        h.set_gemm_tile(-1)
        h.set_tuning(0, 0)


def tune(n: int = 32768, nb: int = 512, out: str = TABLE, verbose: bool = True) -> dict:
    import torch

    import remifa_amd as rf

    h = rf.Handle(load_tuning=False)
    shapes = driver_shapes(n, nb)
    results = {}
    best = {}
    for fam, cand in FAMILIES.items():
        fshapes = [s[1:] for s in shapes if s[0] == fam]
        scores = {}
        for tile in cand["tiles"]:
            for split in cand["splits"]:
                t = sum(_time_gemm(h, rf, torch, M, N, K, tile, split) for M, N, K in fshapes)
                scores[(tile, split)] = t
                if verbose:
                    print(f"{fam:9s} tile {tile} split {split}: {t:8.3f} ms", flush=True)
        (bt, bs), bt_ms = min(scores.items(), key=lambda kv: kv[1])
        best[fam] = {"tile": bt, "split": bs, "ms": bt_ms}
        results[fam] = {f"{t},{s}": round(v, 4) for (t, s), v in scores.items()}
    big = 1 << 40
    rules = [
        # opk, m_lo, m_hi, n_lo, n_hi, tile, split (fp16 and bf16 operands)
        [-1, 1, big, 1, 64, best["narrow_n"]["tile"], best["narrow_n"]["split"]],
        [-1, 1, 64, 65, big, best["narrow_m"]["tile"], best["narrow_m"]["split"]],
        [-1, 65, big, 65, big, best["wide"]["tile"], best["wide"]["split"]],
    ]
    table = {"device": "gfx950", "n": n, "nb": nb, "generated": time.strftime("%Y-%m-%d %H:%M"),
             "rules": rules, "best": best, "candidates_ms": results,
             "note": "rules apply to 16-bit operands; fp32 operands keep the built-in tiles"}
    for r in table["rules"]:
        r[0] = -1
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    return table


def load_rules(path: str = TABLE):
    """rules of a table file, or None"""
    if os.environ.get("REMIFA_TUNING", "").lower() == "off" or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    # fp32 operands only take the 128x128 / 256x32 / 32x256 tiles: restrict
    # the rules to 16-bit operands (kinds 0 and 1)
    rules = []
    for r in t.get("rules", []):
        for opk in (0, 1):
            rules.append([opk] + list(r[1:]))
    return rules


# batched factorization (remifa_lu_nopiv_batched): per size bucket, the
# thread counts its kernel is built for and the persistent workgroups per CU
BATCHED = {64: ([64, 128], [2, 4, 8]), 128: ([128, 256], [1, 2, 4]),
           256: ([256, 512], [1, 2, 3]), 512: ([512], [1, 2])}


def tune_batched(count: int = 4000, verbose: bool = True) -> list:
    """[[bucket, threads, per_cu], ...]: for every size bucket, the fastest
    candidate on `count` matrices with orders uniform over the bucket
    (alpha = 1.2, fp32 storage, TC32 updates)"""
    import numpy as np
    import torch

    import remifa_amd as rf

    dev = torch.device("cuda:0")
    h = rf.Handle(load_tuning=False)
    rng = np.random.default_rng(7)
    rules, lo = [], 16
    for bm, (threads, per_cus) in BATCHED.items():
        sizes = np.sort(rng.integers(lo, bm + 1, count))[::-1].copy()
        lo = bm + 1
        offs = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]])
        a0 = torch.empty(int((sizes.astype(np.int64) ** 2).sum()), dtype=torch.float32, device=dev)
        for i, (n, off) in enumerate(zip(sizes, offs)):
            rf.gen_diagdom(h, a0[off:off + n * n].view(int(n), int(n)).t(), seed=i, alpha=1.2)
        work = torch.empty_like(a0)
        ptrs = torch.tensor(work.data_ptr() + 4 * offs, dtype=torch.int64, device=dev)
        n_t = torch.tensor(sizes, dtype=torch.int32, device=dev)
        info = torch.zeros(count, dtype=torch.int32, device=dev)
        best = None
        for bt in threads:
            for pc in per_cus:
                h.set_batched_tuning([[bm, bt, pc]])
                ms = []
                for _ in range(3):
                    work.copy_(a0)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(h.stream)
                    rf.lu_nopiv_batched(h, rf.TC32, n_t, n_t, ptrs, info, bm)
                    e1.record(h.stream)
                    torch.cuda.synchronize()
                    ms.append(e0.elapsed_time(e1))
                t = min(ms)
                if verbose:
                    print(f"batched bucket {bm:3d} threads {bt:3d} per_cu {pc}: {t:.3f} ms", flush=True)
                if best is None or t < best[0]:
                    best = (t, bt, pc)
        rules.append([bm, best[1], best[2]])
    return rules


def load_batched_rules(path: str = TABLE):
    """batched rules of a table file, or None"""
    if os.environ.get("REMIFA_TUNING", "").lower() == "off" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("batched")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=32768)
    p.add_argument("--nb", type=int, default=512)
    p.add_argument("--out", default=TABLE)
    p.add_argument("--batched-only", action="store_true",
                   help="retune only the batched factorization (keep the GEMM rules)")
    a = p.parse_args()
    if a.batched_only and os.path.exists(a.out):
        with open(a.out) as f:
            t = json.load(f)
    else:
        t = tune(a.n, a.nb, a.out)
    t["batched"] = tune_batched()
    with open(a.out, "w") as f:
        json.dump(t, f, indent=1)
    print(json.dumps({"best": t["best"], "batched": t["batched"]}))


if __name__ == "__main__":
    main()
