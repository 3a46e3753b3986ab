"""The C++ drop-in (include/mpblas_posit.h, include/mplapack_posit.h): a program written
against the reference's API (tests/cpp/dropin_driver.cpp) linked to the HIP library;
its outputs are compared bit for bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from datagen import posit_matrix, uniform_doubles

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "dropin_driver")
ONE, MONE = 0x40000000, 0xC0000000
THREADS = min(16, os.cpu_count() or 1)


def _run(args, tmp_path):
    assert os.path.exists(DRIVER), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _raw(path, n, dtype=np.uint32):
    return np.fromfile(path, dtype=dtype, count=n)


def test_dropin_lu(oracle, tmp_path):
    n, nb = 300, 64
    A = posit_matrix(oracle, 70, n, n)
    A.ravel(order="F").tofile(tmp_path / "a.bin")
    out = _run(["lu", str(tmp_path / "a.bin"), str(n), str(nb), str(tmp_path / "lu")], tmp_path)
    F = A.copy(order="F")
    _, ipiv, info = oracle.rgetrf(F, nb=nb, nthreads=THREADS)
    assert out.split()[:3] == ["info", str(info), "0"]
    assert np.array_equal(_raw(tmp_path / "lu_factor.bin", n * n), F.ravel(order="F"))
    assert np.array_equal(_raw(tmp_path / "lu_ipiv.bin", n, np.int32), ipiv)
    b = np.zeros(n, np.uint32)
    oracle.rgemv("N", n, n, ONE, A, np.full(n, ONE, np.uint32), 1, 0, b, 1)
    B = np.asfortranarray(b.reshape(n, 1))
    oracle.rgetrs("N", F, ipiv, B)
    assert np.array_equal(_raw(tmp_path / "lu_x.bin", n), B[:, 0])


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dropin_cholesky(oracle, tmp_path, uplo):
    n = 200
    M = uniform_doubles(71, n * n).reshape(n, n)
    S = M @ M.T / n + np.eye(n)
    A = np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))
    A.ravel(order="F").tofile(tmp_path / "a.bin")
    out = _run(["cho", str(tmp_path / "a.bin"), str(n), uplo, str(tmp_path / "ch")], tmp_path)
    F = A.copy(order="F")
    assert oracle.rpotrf(uplo, F, nb=0, nthreads=THREADS) == (0, 0)
    assert out.split()[:3] == ["info", "0", "0"]
    assert np.array_equal(_raw(tmp_path / "ch_factor.bin", n * n), F.ravel(order="F"))
    b = np.zeros(n, np.uint32)
    oracle.rgemv("N", n, n, ONE, A, np.full(n, ONE, np.uint32), 1, 0, b, 1)
    B = np.asfortranarray(b.reshape(n, 1))
    oracle.rpotrs(uplo, F, B)
    assert np.array_equal(_raw(tmp_path / "ch_x.bin", n), B[:, 0])


def test_dropin_rgemmg_and_scalars(oracle, tmp_path):
    n, k = 160, 48
    A = posit_matrix(oracle, 72, n, n)
    A.ravel(order="F").tofile(tmp_path / "a.bin")
    out = _run(["gemm", str(tmp_path / "a.bin"), str(n), str(k), str(tmp_path / "g")], tmp_path)
    C = A.copy(order="F")
    oracle.rgemm("N", "N", n, n, k, MONE, A, A, ONE, C, nthreads=THREADS, lda=n, ldb=n)
    assert np.array_equal(_raw(tmp_path / "g_c.bin", n * n), C.ravel(order="F"))
    # posit32 scalar operators: ((0.75 * -2.5) + 1) / -2.5 and sqrt(2)
    x, y = oracle.from_double(0.75), oracle.from_double(-2.5)
    z = oracle.div(oracle.add(oracle.mul(x, y), ONE), y)
    s = oracle.sqrt(oracle.from_double(2.0))
    words = out.split()
    assert words[0] == "scalar" and int(words[1], 16) == z and int(words[2], 16) == s
