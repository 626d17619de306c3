"""remifa_getrs_nopiv_noalloc (the solve whose scratch is a static device
array, for callers that cannot allocate) against remifa_getrs_nopiv, through
the C ABI: same kernels, so the solutions must be bit-identical. Two calls
in a row, the second with a smaller order (the array keeps the first call's
contents), both triangles orders (plain and transposed), fp16 and fp32
factors; orders above the array's maximum are rejected with
REMIFA_ERR_WORKSPACE and a message."""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from helpers import to_dev, to_np

pytestmark = pytest.mark.gpu


def _solve(lib, h, fn, lu, b):
    n = lu.shape[0]
    lu_t = 1 if lu.dtype.itemsize == 2 else 0
    return getattr(lib, fn)(h, lu_t, 0, n, b.shape[1], C.c_void_p(lu.data_ptr()), lu.stride(1),
                            C.c_void_p(b.data_ptr()), b.stride(1))


@pytest.mark.parametrize("storage", ["f16", "f32"])
def test_noalloc_solve_matches_getrs(gpu, storage):
    import torch
    import remifa_amd as rf
    from remifa_amd._lib import lib
    h = rf.Handle()
    L = lib()
    for trans in (0, 1):
        for n, nrhs, seed in ((1000, 3, 1), (300, 1, 2), (129, 17, 3)):
            a32 = orc.gen_diagdom_alpha(n, seed, 1.3)
            if storage == "f16":
                lu_raw = orc.lu_nopiv("ll", orc.cfg(W=orc.F16, I=orc.TC32, O=orc.TC32,
                                                    OP=orc.F32, ib=32),
                                      orc.f32_to_f16_bits(a32), 64)
            else:
                lu_raw = orc.lu_nopiv("ll", orc.cfg(), a32, 64)
            dlu = to_dev(lu_raw, gpu, ld=n + 2)
            b = np.random.default_rng(seed).uniform(-1, 1, (n, nrhs)).astype(np.float32)
            b1 = to_dev(np.asfortranarray(b), gpu, ld=n + 1)
            b2 = to_dev(np.asfortranarray(b), gpu, ld=n + 1)
            lu_t = 1 if storage == "f16" else 0
            args = (h._h, lu_t, trans, n, nrhs, C.c_void_p(dlu.data_ptr()), dlu.stride(1))
            assert L.remifa_getrs_nopiv(*args, C.c_void_p(b1.data_ptr()), b1.stride(1)) == 0
            assert L.remifa_getrs_nopiv_noalloc(*args, C.c_void_p(b2.data_ptr()),
                                                b2.stride(1)) == 0
            torch.cuda.synchronize()
            x1, x2 = to_np(b1), to_np(b2)
            assert np.array_equal(x1.view(np.uint32), x2.view(np.uint32)), (trans, n, nrhs)
            # and both solve the system as the oracle's FP64 substitution does
            lu = orc.f16_bits_to_f32(lu_raw) if storage == "f16" else lu_raw
            xr = orc.getrs_f64(lu.astype(np.float64), b.astype(np.float64), trans=bool(trans))
            assert np.max(np.abs(x2 - xr)) <= 1e-4 * max(np.max(np.abs(xr)), 1.0)


def test_noalloc_solve_rejects_orders_above_its_scratch(gpu):
    import torch
    import remifa_amd as rf
    from remifa_amd._lib import ERR_WORKSPACE, lib
    h = rf.Handle()
    lu = torch.zeros(1, dtype=torch.float32, device=gpu)
    b = torch.zeros(1, dtype=torch.float32, device=gpu)
    n = 32769
    st = lib().remifa_getrs_nopiv_noalloc(h._h, 0, 0, n, 1, C.c_void_p(lu.data_ptr()), n,
                                          C.c_void_p(b.data_ptr()), n)
    assert st == ERR_WORKSPACE
    assert b"32768" in lib().remifa_last_error()
    assert lib().remifa_getrs_nopiv_noalloc(None, 0, 0, 4, 1, None, 4, None, 4) == -1
    assert lib().remifa_getrs_nopiv_noalloc(h._h, 0, 0, -1, 1, None, 4, None, 4) == -4
