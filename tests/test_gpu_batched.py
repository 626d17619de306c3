"""Batched no-pivot LU + solve (remifa_lu_nopiv_batched): every matrix of a
ragged batch against the oracle's restatement of the RL driver with
nb = ib = 32 and TC32 updates (src/lu_nopiv.hxx:53-208), against the FP64
LU, order-independence of the persistent queue, the batched solve, and
zero-pivot reporting."""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity

pytestmark = pytest.mark.gpu


def _batch(sizes, seed, alpha=1.2):
    mats = [orc.gen_diagdom_alpha(int(n), seed + i, alpha) for i, n in enumerate(sizes)]
    return mats


def _pack(mats, gpu, pad=3):
    """one device buffer, matrix i column-major at offset off[i] with ld n+pad"""
    import torch
    lds = [m.shape[0] + pad for m in mats]
    sizes = [ld * m.shape[1] for ld, m in zip(lds, mats)]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = torch.zeros(int(sum(sizes)) + 64, dtype=torch.float32)
    for m, ld, off in zip(mats, lds, offs):
        n = m.shape[0]
        blk = np.zeros((n, ld), np.float32)  # column j at row j of blk
        blk[:, :n] = m.T
        buf[off:off + ld * n] = torch.from_numpy(blk.reshape(-1))
    buf = buf.to(gpu)
    ptrs = torch.tensor([buf.data_ptr() + 4 * int(o) for o in offs], dtype=torch.int64, device=gpu)
    n_t = torch.tensor([m.shape[0] for m in mats], dtype=torch.int32, device=gpu)
    ld_t = torch.tensor(lds, dtype=torch.int32, device=gpu)
    return buf, ptrs, n_t, ld_t, offs, lds


def _unpack(buf, mats, offs, lds):
    host = buf.cpu().numpy()
    out = []
    for m, off, ld in zip(mats, offs, lds):
        n = m.shape[0]
        out.append(np.asfortranarray(host[off:off + ld * n].reshape(n, ld)[:, :n].T))
    return out


def _run(handle, gpu, mats, ctype=None):
    import torch
    import remifa_amd as rf
    buf, ptrs, n_t, ld_t, offs, lds = _pack(mats, gpu)
    info = torch.full((len(mats),), -7, dtype=torch.int32, device=gpu)
    rf.lu_nopiv_batched(handle, rf.TC32 if ctype is None else ctype, n_t, ld_t, ptrs, info,
                        int(max(m.shape[0] for m in mats)))
    torch.cuda.synchronize()
    return _unpack(buf, mats, offs, lds), info.cpu().numpy()


def test_batched_lu_matches_oracle(handle, gpu):
    rng = np.random.default_rng(1)
    sizes = list(rng.integers(16, 513, 24)) + [2, 3, 31, 32, 33, 64, 65, 512]
    mats = _batch(sizes, 100)
    got, info = _run(handle, gpu, mats)
    assert np.all(info == 0)
    c = orc.cfg(W=orc.F32, I=orc.TC32, O=orc.TC32, ib=32)
    for a, lu in zip(mats, got):
        o = orc.lu_nopiv("rl", c, a, 32)
        # same rounding points as the oracle; only the order of the fp32
        # sums inside the matrix-core update differs: entrywise parity
        assert_parity(lu, o, f"batched n={a.shape[0]}")
        f64 = orc.lu_f64(a)
        e_o = np.max(np.abs(o - f64)) / np.max(np.abs(f64))
        e_g = np.max(np.abs(lu - f64)) / np.max(np.abs(f64))
        assert e_g <= 8 * max(e_o, 2.0**-24), (a.shape, e_g, e_o)


def test_batched_bf16_update(handle, gpu):
    import remifa_amd as rf
    mats = _batch([40, 200, 333], 7)
    got, info = _run(handle, gpu, mats, ctype=rf.BF16_TC32)
    assert np.all(info == 0)
    for a, lu in zip(mats, got):
        f64 = orc.lu_f64(a)
        err = np.max(np.abs(lu - f64)) / np.max(np.abs(f64))
        assert err < 64 * 2.0**-8


def test_batched_result_independent_of_batch(handle, gpu):
    """a matrix's factors do not depend on its neighbours or position"""
    mats = _batch([100, 300, 17, 512, 64], 55)
    got1, _ = _run(handle, gpu, mats)
    order = [3, 0, 4, 2, 1]
    got2, _ = _run(handle, gpu, [mats[i] for i in order] + _batch([250, 90], 9))
    for j, i in enumerate(order):
        assert np.array_equal(got1[i].view(np.uint32), got2[j].view(np.uint32))


@pytest.mark.parametrize("bucket,threads,per_cu", [(64, 64, 4), (64, 128, 2), (128, 128, 4),
                                                   (128, 256, 1), (256, 256, 2), (256, 512, 1),
                                                   (512, 512, 2)])
def test_batched_every_bucket_kernel_matches_oracle(handle, gpu, bucket, threads, per_cu):
    """every kernel the tuning table can pick (size bucket x threads x
    workgroups per CU): ragged orders inside the bucket, against the oracle;
    the factors do not depend on the shape (same operations per entry)"""
    rng = np.random.default_rng(bucket + threads)
    lo = {64: 2, 128: 65, 256: 129, 512: 257}[bucket]
    sizes = list(rng.integers(lo, bucket + 1, 5)) + [bucket]
    mats = _batch(sizes, 300 + bucket)
    try:
        handle.set_batched_tuning([[bucket, threads, per_cu]])
        got, info = _run(handle, gpu, mats)
        handle.set_batched_tuning([[bucket, bucket, 1]])
        ref, _ = _run(handle, gpu, mats)
    finally:
        handle.set_batched_tuning([])
    assert np.all(info == 0)
    c = orc.cfg(W=orc.F32, I=orc.TC32, O=orc.TC32, ib=32)
    for a, lu, r in zip(mats, got, ref):
        assert_parity(lu, orc.lu_nopiv("rl", c, a, 32), f"bucket {bucket} n={a.shape[0]}")
        assert np.array_equal(lu.view(np.uint32), r.view(np.uint32))


def test_batched_tuning_rejects_bad_rules(handle, gpu):
    import remifa_amd as rf
    for bad in ([[96, 128, 1]], [[128, 64, 1]], [[512, 256, 1]], [[256, 256, 0]],
                [[64, 64, 17]]):
        with pytest.raises(rf.RemifaError):
            handle.set_batched_tuning(bad)
    handle.set_batched_tuning([[512, 512, 2], [64, 128, 8]])
    handle.set_batched_tuning([])


def test_batched_solve(handle, gpu):
    import torch
    import remifa_amd as rf
    rng = np.random.default_rng(3)
    sizes = list(rng.integers(16, 513, 12)) + [2, 33]
    mats = _batch(sizes, 300)
    buf, ptrs, n_t, ld_t, offs, lds = _pack(mats, gpu)
    info = torch.zeros(len(mats), dtype=torch.int32, device=gpu)
    rf.lu_nopiv_batched(handle, rf.TC32, n_t, ld_t, ptrs, info, 512)
    bs = [rng.uniform(-1, 1, m.shape[0]).astype(np.float32) for m in mats]
    bbuf = [torch.from_numpy(b).to(gpu) for b in bs]
    bptrs = torch.tensor([t.data_ptr() for t in bbuf], dtype=torch.int64, device=gpu)
    rf.getrs_nopiv_batched(handle, n_t, ld_t, ptrs, bptrs, 512)
    torch.cuda.synchronize()
    facs = _unpack(buf, mats, offs, lds)
    c = orc.cfg(W=orc.F32, I=orc.TC32, O=orc.TC32, ib=32)
    for a, lu, b, xb in zip(mats, facs, bs, bbuf):
        x = xb.cpu().numpy()
        want = orc.getrs_f32(lu, b)[:, 0]
        assert np.max(np.abs(x - want)) <= 1e-5 * max(np.max(np.abs(want)), 1.0)
        # the factors carry the fp16 operand rounding of the updates: the
        # solve is as good as the oracle's factors allow
        xo = orc.getrs_f32(orc.lu_nopiv("rl", c, a, 32), b)[:, 0]
        be_o = orc.backward_error(a, b.astype(np.float64), xo.astype(np.float64))
        be = orc.backward_error(a, b.astype(np.float64), x.astype(np.float64))
        assert be <= 4 * be_o + 1e-7


def test_batched_zero_pivot_reported(handle, gpu):
    a = orc.gen_diagdom_alpha(70, 4, 1.2)
    a[40, :] = 0.0  # row 40 zero -> u_40,40 = 0 at column 41
    one = np.zeros((1, 1), np.float32)  # 1 x 1 zero matrix: pivot 1
    mats = [orc.gen_diagdom_alpha(50, 5, 1.2), a, one]
    _, info = _run(handle, gpu, mats)
    assert info[0] == 0 and info[1] == 41 and info[2] == 1


def test_batched_lu_and_solve_in_hip_graph(handle, gpu):
    """remifa_lu_nopiv_batched + remifa_getrs_nopiv_batched captured in one
    HIP graph (the queue counter is reset by a kernel inside the graph, no
    memset node), replayed four times with eager work and synchronisations
    between the replays: every replay's factors and solutions equal the
    eager call's bit for bit, and every matrix matches the oracle."""
    import torch
    import remifa_amd as rf
    rng = np.random.default_rng(21)
    sizes = list(rng.integers(16, 513, 40)) + [16, 512]
    sizes = sorted(sizes, reverse=True)  # largest first, as the bench orders them
    mats = _batch(sizes, 700)
    buf, ptrs, n_t, ld_t, offs, lds = _pack(mats, gpu)
    a0 = buf.clone()
    bs = [rng.uniform(-1, 1, m.shape[0]).astype(np.float32) for m in mats]
    boffs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    b0 = torch.from_numpy(np.concatenate(bs)).to(gpu)
    bw = b0.clone()
    bptrs = torch.tensor([bw.data_ptr() + 4 * int(o) for o in boffs], dtype=torch.int64,
                         device=gpu)
    info = torch.full((len(mats),), -7, dtype=torch.int32, device=gpu)

    def run(h):
        rf.lu_nopiv_batched(h, rf.TC32, n_t, ld_t, ptrs, info, 512)
        rf.getrs_nopiv_batched(h, n_t, ld_t, ptrs, bptrs, 512)

    run(handle)  # eager (also sizes the workspace before the capture)
    torch.cuda.synchronize()
    eager_f, eager_x, eager_i = buf.clone(), bw.clone(), info.clone()
    gs = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    old = handle.stream
    try:
        handle.set_stream(gs)
        with torch.cuda.graph(graph, stream=gs):
            run(handle)
    finally:
        handle.set_stream(old)
    for rep in range(4):
        buf.copy_(a0)
        bw.copy_(b0)
        info.fill_(-7)
        torch.cuda.synchronize()
        with torch.cuda.stream(gs):
            graph.replay()
        torch.cuda.synchronize()
        # eager work between the replays
        _ = (buf[:1000] * 2.0).sum().item()
        assert torch.equal(info, eager_i), rep
        assert torch.equal(buf.view(torch.int32), eager_f.view(torch.int32)), rep
        assert torch.equal(bw.view(torch.int32), eager_x.view(torch.int32)), rep
    got = _unpack(buf, mats, offs, lds)
    c = orc.cfg(W=orc.F32, I=orc.TC32, O=orc.TC32, ib=32)
    xs = bw.cpu().numpy()
    for i, (a, lu) in enumerate(zip(mats, got)):
        assert_parity(lu, orc.lu_nopiv("rl", c, a, 32), f"graph batched n={a.shape[0]}")
        n = a.shape[0]
        x = xs[boffs[i]:boffs[i] + n]
        want = orc.getrs_f32(lu, bs[i])[:, 0]
        assert np.max(np.abs(x - want)) <= 1e-5 * max(np.max(np.abs(want)), 1.0)
