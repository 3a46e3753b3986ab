"""Batched posit32 GEMM (one persistent launch, autotuned, hipGraph replay): every problem
must equal the plain per-call result and the oracle bit for bit."""
import os

import numpy as np
import pytest

from datagen import posit_matrix

pytestmark = pytest.mark.gpu

ONE, MONE = 0x40000000, 0xC0000000
THREADS = min(16, os.cpu_count() or 1)


def _make_batch(pga, oracle, rng, count, lo=16, hi=256, step=16):
    probs, host = [], []
    for i in range(count):
        m, n, k = (int(x) for x in rng.integers(lo // step, hi // step + 1, 3) * step)
        A = posit_matrix(oracle, 1000 + 3 * i, m, k)
        B = posit_matrix(oracle, 1001 + 3 * i, k, n)
        C = posit_matrix(oracle, 1002 + 3 * i, m, n)
        dA, dB, dC = pga.DeviceArray.from_host(A), pga.DeviceArray.from_host(B), pga.DeviceArray.from_host(C)
        alpha = [ONE, MONE, 0x3A000000][i % 3]
        beta = [ONE, 0, 0x44000000][i % 3]
        probs.append(dict(m=m, n=n, k=k, alpha=alpha, beta=beta, a=dA, lda=m, b=dB, ldb=k, c=dC, ldc=m))
        host.append((A, B, C))
    return probs, host


@pytest.mark.parametrize("tile", [0, 16, 32, 64])
def test_batched_equals_plain_and_oracle(pga, oracle, tile):
    rng = np.random.default_rng(5 + tile)
    probs, host = _make_batch(pga, oracle, rng, 40)
    plan = pga.BatchPlan("N", "N", probs, tile=tile)
    plan.execute()
    pga.synchronize()
    for p, (A, B, C) in zip(probs, host):
        got = p["c"].to_host()
        want = C.copy(order="F")
        oracle.rgemm("N", "N", p["m"], p["n"], p["k"], p["alpha"], A, B, p["beta"], want, nthreads=THREADS)
        assert np.array_equal(got, want)
        plain = pga.rgemm("N", "N", p["alpha"], A, B, p["beta"], C)
        assert np.array_equal(got, plain)
    if tile == 0:
        info = plan.info()
        assert info["tile"] in (16, 32, 64) and info["grid"] > 0 and info["schedule"] in (0, 1, 2)


def test_batched_graph_replays(pga, oracle):
    """Capture once, replay several times with eager work in between; C accumulates
    C = A*B + C per replay, checked against plain calls after every replay."""
    rng = np.random.default_rng(77)
    probs, host = _make_batch(pga, oracle, rng, 24)
    for p in probs:
        p["alpha"], p["beta"] = ONE, ONE
    plan = pga.BatchPlan("N", "N", probs, tile=32)
    plan.capture_graph()
    expect = [C.copy(order="F") for (_, _, C) in host]
    for rep in range(3):
        plan.launch_graph()
        pga.synchronize()
        for i, (p, (A, B, _)) in enumerate(zip(probs, host)):
            expect[i] = pga.rgemm("N", "N", ONE, A, B, ONE, expect[i])  # eager work between replays
            assert np.array_equal(p["c"].to_host(), expect[i]), (rep, i)


@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "T")])
def test_batched_transposes(pga, oracle, ta, tb):
    rng = np.random.default_rng(9)
    probs, host = [], []
    for i in range(12):
        m, n, k = (int(x) for x in rng.integers(1, 9, 3) * 16)
        A = posit_matrix(oracle, 2000 + i, k if ta == "T" else m, m if ta == "T" else k)
        B = posit_matrix(oracle, 2100 + i, n if tb == "T" else k, k if tb == "T" else n)
        C = posit_matrix(oracle, 2200 + i, m, n)
        dA, dB, dC = pga.DeviceArray.from_host(A), pga.DeviceArray.from_host(B), pga.DeviceArray.from_host(C)
        probs.append(dict(m=m, n=n, k=k, alpha=MONE, beta=ONE, a=dA, lda=A.shape[0], b=dB, ldb=B.shape[0], c=dC,
                          ldc=m))
        host.append((A, B, C))
    plan = pga.BatchPlan(ta, tb, probs, tile=0)
    plan.execute()
    pga.synchronize()
    for p, (A, B, C) in zip(probs, host):
        want = C.copy(order="F")
        oracle.rgemm(ta, tb, p["m"], p["n"], p["k"], MONE, A, B, ONE, want, nthreads=THREADS)
        assert np.array_equal(p["c"].to_host(), want)
