"""pga_batch_gemm_noalloc (the batched GEMM that allocates no device memory) against the
plan-based batched GEMM (pga_batch_plan_create + execute) on the same problems, through
the C ABI: two calls in a row, the second with fewer problems and other transposes,
every C bit for bit; a batch larger than the fixed scratch is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ONE, MONE = 0x40000000, 0xC0000000


def _batch(pga, rng, count, ta, tb, seed):
    """count problems with m, n, k multiples of 16 in [16, 256] in three device buffers
    (A, B and two copies of C), alternating (alpha, beta) = (-1, 1) and (1, 0)."""
    dims = rng.integers(1, 17, size=(count, 3)) * 16
    offs = np.zeros((count + 1, 3), np.int64)
    for i, (m, n, k) in enumerate(dims):
        offs[i + 1] = offs[i] + [m * k, k * n, m * n]
    dA = pga.DeviceArray(int(offs[-1, 0]), np.uint32, "C")
    dB = pga.DeviceArray(int(offs[-1, 1]), np.uint32, "C")
    dC1 = pga.DeviceArray(int(offs[-1, 2]), np.uint32, "C")
    dC2 = pga.DeviceArray(int(offs[-1, 2]), np.uint32, "C")
    pga.fill_uniform(dA, seed)
    pga.fill_uniform(dB, seed + 1)
    pga.fill_uniform(dC1, seed + 2)
    dC2.copy_from(dC1)
    probs = ([], [])
    for i, (m, n, k) in enumerate(dims):
        m, n, k = int(m), int(n), int(k)
        alpha, beta = (MONE, ONE) if i % 2 == 0 else (ONE, 0)
        base = {"m": m, "n": n, "k": k, "alpha": alpha, "beta": beta,
                "a": dA.ptr + 4 * int(offs[i, 0]), "lda": k if ta else m,
                "b": dB.ptr + 4 * int(offs[i, 1]), "ldb": n if tb else k, "ldc": m}
        for dC, lst in ((dC1, probs[0]), (dC2, probs[1])):
            lst.append(dict(base, c=dC.ptr + 4 * int(offs[i, 2])))
    return probs, (dA, dB, dC1, dC2)


def test_noalloc_equals_plan(pga):
    rng = np.random.default_rng(2024)
    for count, ta, tb, seed in ((600, "N", "N", 10), (75, "T", "N", 20)):
        (p_plan, p_na), bufs = _batch(pga, rng, count, ta == "T", tb == "T", seed)
        plan = pga.BatchPlan(ta, tb, p_plan, tile=64)
        plan.execute()
        pga.batch_gemm_noalloc(ta, tb, p_na)
        pga.synchronize()
        want, got = bufs[2].to_host(), bufs[3].to_host()
        assert np.array_equal(got, want), (count, int((got != want).sum()))
        del plan


def test_noalloc_refuses_batch_beyond_scratch(pga):
    cnt = pga.NOALLOC_MAX_PROBLEMS + 1
    ints = (ctypes.c_int * cnt)(*([16] * cnt))
    u32 = (ctypes.c_uint32 * cnt)(*([ONE] * cnt))
    ptrs = (ctypes.c_void_p * cnt)()
    st = pga._lib.pga_batch_gemm_noalloc(b"N", b"N", cnt, ints, ints, ints, u32, ptrs, ints, ptrs, ints, u32,
                                         ptrs, ints, 0, 0, None)
    assert st == 7  # PGA_ERROR_EXCEEDS_SCRATCH, before any device work
