"""Full-size parity against the oracle: the device results of the configurations at their
real sizes, compared with SHA-256 digests of the oracle's results on the same seeded
inputs (tests/golden/fullsize_hashes.json, written by tests/golden/gen_fullsize_hashes.py,
which documents the workloads).

  * the four trailing-update GEMMs of a blocked LU at n = 16384 (M = N in {16384, 12288,
    8192, 4096}, K = 256, alpha = -1, beta = 1): every element of C;
  * Rgetrf (nb = 256) at n = 8192 with Rgetrs of b = A*1: factors, pivots, info, solution;
  * lower Rpotrf (nb = 64) at n = 8192 with Rpotrs of b = S*1: factor, info, solution.

The input digests are checked first, so that a generator mismatch is told apart from a
wrong result; a wrong result names the 1024-column blocks that differ."""
import json
import os

import numpy as np
import pytest

from datagen import uniform01
from gen_fullsize_hashes import BLOCK, digest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "fullsize_hashes.json")
ONE, MONE = 0x40000000, 0xC0000000


def _fixture(section):
    if not os.path.exists(FIXTURE):
        pytest.fail(f"missing fixture {FIXTURE}")
    with open(FIXTURE) as f:
        fx = json.load(f)
    if section not in fx or not fx[section]:
        pytest.fail(f"fixture has no '{section}' entry")
    return fx[section]


def _same(got, ent, what):
    """Whole-array digest against the fixture; on a mismatch name the differing blocks."""
    if digest(got) == ent["sha256"]:
        return
    if got.ndim == 2:
        parts = [digest(got[:, j:j + BLOCK]) for j in range(0, got.shape[1], BLOCK)]
    else:
        parts = [digest(got[j:j + BLOCK]) for j in range(0, got.shape[0], BLOCK)]
    bad = [i for i, (a, b) in enumerate(zip(parts, ent["blocks"])) if a != b]
    pytest.fail(f"{what}: differs from the oracle in column blocks {bad} (of {len(parts)}, {BLOCK} wide)")


@pytest.mark.parametrize("size", [16384, 12288, 8192, 4096])
def test_trailing_gemm_full_size(pga, size):
    ent = _fixture("gemm").get(str(size))
    assert ent is not None, size
    k = ent["k"]
    sa, sb, sc = ent["seeds"]
    dA = pga.DeviceArray((size, k))
    dB = pga.DeviceArray((k, size))
    dC = pga.DeviceArray((size, size))
    pga.fill_uniform(dA, sa)
    pga.fill_uniform(dB, sb)
    pga.fill_uniform(dC, sc)
    assert digest(dA.to_host()) == ent["a"] and digest(dB.to_host()) == ent["b"]
    assert digest(dC.to_host()) == ent["c_in"]
    pga.rgemm_device("N", "N", size, size, k, MONE, dA, size, dB, k, ONE, dC, size)
    pga.synchronize()
    _same(dC.to_host(), ent["c_out"], f"C of the {size} x {size} x {k} GEMM")


def test_lu_8192_full_size(pga):
    ent = _fixture("lu")
    n, nb = ent["n"], ent["nb"]
    dA0 = pga.DeviceArray((n, n))
    pga.fill_uniform(dA0, ent["seed"])
    assert digest(dA0.to_host()) == ent["a"]
    ones = pga.DeviceArray.from_host(np.full(n, ONE, np.uint32))
    db = pga.DeviceArray(n, np.uint32, "C")
    pga.gemv_device("N", n, n, ONE, dA0, n, ones, 1, 0, db, 1)
    pga.synchronize()
    assert digest(db.to_host()) == ent["b"]
    dp = pga.DeviceArray(n, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(n, n, dA0, n, dp, di, nb)
    pga.synchronize()
    assert int(di.to_host()[0]) == ent["info"]
    assert digest(dp.to_host().astype(np.int32)) == ent["ipiv"]
    _same(dA0.to_host(), ent["factor"], "LU factors")
    dx = pga.DeviceArray((n, 1))
    dx.copy_from(db)
    pga.getrs_device("N", n, 1, dA0, n, dp, dx, n)
    pga.synchronize()
    assert digest(dx.to_host()[:, 0]) == ent["x"]


def test_cholesky_8192_full_size(pga):
    ent = _fixture("chol")
    n, nb = ent["n"], ent["nb"]
    # S = M M^T / n + I with M on a 2^-20 grid: M M^T is exact in FP64 in any order
    u = uniform01(ent["seed"], n * n)
    M = (np.floor(u * 2.0 ** 21) - 2.0 ** 20).reshape(n, n)
    del u
    S = M @ M.T
    del M
    S *= 2.0 ** -40
    S /= n
    S[np.diag_indices(n)] += 1.0
    dS = pga.DeviceArray.from_host(np.asfortranarray(S))
    del S
    dA = pga.DeviceArray((n, n))
    pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, dA.ptr, n * n, None), "from_f64")
    pga.synchronize()
    dS.free()
    assert digest(dA.to_host()) == ent["a"]
    ones = pga.DeviceArray.from_host(np.full(n, ONE, np.uint32))
    db = pga.DeviceArray(n, np.uint32, "C")
    pga.gemv_device("N", n, n, ONE, dA, n, ones, 1, 0, db, 1)
    pga.synchronize()
    assert digest(db.to_host()) == ent["b"]
    di = pga.DeviceArray(1, np.int32, "C")
    pga.potrf_device(ent["uplo"], n, dA, n, di, nb)
    pga.synchronize()
    assert int(di.to_host()[0]) == ent["info"]
    _same(dA.to_host(), ent["factor"], "Cholesky factor")
    dx = pga.DeviceArray((n, 1))
    dx.copy_from(db)
    pga.potrs_device(ent["uplo"], n, 1, dA, n, dx, n)
    pga.synchronize()
    assert digest(dx.to_host()[:, 0]) == ent["x"]
