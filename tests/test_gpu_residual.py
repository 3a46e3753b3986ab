"""The FP64 backward error printed beside LU / Cholesky results
(pga_backward_error_device) against the same figure formed on the host: A, x and b
decoded exactly by the oracle, each row of A x - b summed with math.fsum (correctly
rounded), norms as maxima.  The device sums each row in FP64 in ascending column order,
so the two agree to the rounding of an FP64 dot product, not bit for bit."""
import math
import os

import numpy as np
import pytest

from datagen import posit_matrix

pytestmark = pytest.mark.gpu

ONE = 0x40000000
THREADS = min(16, os.cpu_count() or 1)


def _host_berr(oracle, A, x, b):
    Ad, xd, bd = oracle.to_double_v(A), oracle.to_double_v(x), oracle.to_double_v(b)
    n = Ad.shape[0]
    r = max(abs(math.fsum(list(Ad[i] * xd) + [-bd[i]])) for i in range(n))
    # the products a_ij * x_j of posits are exact in FP64 (at most 28 x 28 significant bits)
    anorm = max(math.fsum(np.abs(Ad[i])) for i in range(n))
    xnorm = np.abs(xd).max()
    return r / (anorm * xnorm)


@pytest.mark.parametrize("case", ["random", "solution"])
def test_backward_error_matches_host(pga, oracle, case):
    n = 257
    A = posit_matrix(oracle, 900, n, n)
    b = posit_matrix(oracle, 901, n, 1)[:, 0].copy()
    if case == "random":  # a large residual: no cancellation in A x - b
        x = posit_matrix(oracle, 902, n, 1)[:, 0].copy()
        rtol = 1e-12
    else:  # x from the LU solve: the residual is the small difference of large sums
        F = A.copy(order="F")
        _, piv, info = oracle.rgetrf(F, nb=64, nthreads=THREADS)
        assert info == 0
        X = np.asfortranarray(b.reshape(n, 1))
        oracle.rgetrs("N", F, piv, X)
        x = X[:, 0].copy()
        rtol = 1e-6  # FP64 summation order of a residual ~1e-7 of |A||x|
    want = _host_berr(oracle, A, x, b)
    dA = pga.DeviceArray.from_host(A)
    dx = pga.DeviceArray.from_host(x)
    db = pga.DeviceArray.from_host(b)
    got = pga.backward_error_device(n, dA, n, dx, db)
    assert want > 0
    assert abs(got - want) <= rtol * want, (got, want)
