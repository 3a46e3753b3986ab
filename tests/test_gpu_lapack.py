"""GPU parity of the posit32 BLAS-2/3 and LU / Cholesky routines (through the C ABI)
against the oracle's restatements of the reference routines.  Bit-exact: factors,
pivots, info codes and solutions."""
import os

import numpy as np
import pytest

from datagen import posit_matrix, uniform_doubles

pytestmark = pytest.mark.gpu

ONE, MONE, NAR = 0x40000000, 0xC0000000, 0x80000000
THREADS = min(16, os.cpu_count() or 1)


def _same(got, want):
    bad = np.argwhere(np.asarray(got) != np.asarray(want))
    assert bad.size == 0, f"{len(bad)} mismatches, first {bad[:4].tolist()}"


def _tri(oracle, seed, n, upper, unit_scale=True):
    """Well-conditioned triangular posit matrix (dominant diagonal)."""
    T = posit_matrix(oracle, seed, n, n, -0.5, 0.5)
    d = oracle.from_double_v(np.linspace(1.0, 2.0, n))
    for i in range(n):
        T[i, i] = d[i]
    return T


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("transa", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("alpha", [ONE, 0x3A000000, 0])
def test_trsm(pga, oracle, side, uplo, transa, diag, alpha):
    m, n = 37, 23
    na = m if side == "L" else n
    A = _tri(oracle, 5, na, uplo == "U")
    B = posit_matrix(oracle, 6, m, n)
    B[3, :] = 0  # zero right-hand-side entries exercise the reference's skips
    want = B.copy(order="F")
    assert oracle.rtrsm(side, uplo, transa, diag, m, n, alpha, A, want) == 0
    got = B.copy(order="F")
    pga.Rtrsm(side, uplo, transa, diag, m, n, alpha, A, na, got, m)
    _same(got, want)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("kind", ["huge", "tiny", "nar", "cancel"])
def test_trsm_diagonal_block_classes(pga, oracle, uplo, diag, kind):
    """Left non-transposed solves of up to 64 rows run one wave per right-hand side in
    FP64 and redo a column at the pattern level when a value leaves the fast range or is
    NaR: scaled far up or down, a NaR inside the triangle, and cancelling sums."""
    m, n = 48, 40
    rng = np.random.default_rng({"huge": 81, "tiny": 82, "nar": 83, "cancel": 84}[kind])
    if kind == "cancel":
        Ad = rng.choice([-1.0, -0.5, 0.5, 1.0], (m, m))
        Bd = rng.choice([-1.0, -0.5, 0.5, 1.0, 0.0], (m, n))
    else:
        Ad = rng.uniform(-1, 1, (m, m))
        Bd = rng.uniform(-1, 1, (m, n))
        if kind == "huge":
            Ad *= 2.0 ** 50
            Bd *= 2.0 ** 55
        elif kind == "tiny":
            Ad *= 2.0 ** -50
            Bd *= 2.0 ** -58
    Ad += np.eye(m) * (2.0 if kind != "huge" else 2.0 ** 51)
    Ad = np.triu(Ad) if uplo == "U" else np.tril(Ad)
    A = np.asfortranarray(oracle.from_double_v(Ad.ravel(order="F")).reshape(m, m, order="F"))
    B = np.asfortranarray(oracle.from_double_v(Bd.ravel(order="F")).reshape(m, n, order="F"))
    if kind == "nar":
        if uplo == "L":
            A[30, 12] = NAR
        else:
            A[12, 30] = NAR
        B[5, 7] = NAR
    want = B.copy(order="F")
    assert oracle.rtrsm("L", uplo, "N", diag, m, n, ONE, A, want) == 0
    got = B.copy(order="F")
    pga.Rtrsm("L", uplo, "N", diag, m, n, ONE, A, m, got, m)
    _same(got, want)


@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("special", [False, True])
def test_trsm_left_lower_tall(pga, oracle, diag, special):
    """Left lower solves with many rows (solved in row blocks with GEMM updates)."""
    m, n = 150, 37
    A = _tri(oracle, 72, m, False)
    B = posit_matrix(oracle, 73, m, n)
    B[40, :] = 0  # zero b_k skip their step
    B[3, ::2] = 0
    if special:  # NaR in L under skipped and unskipped multipliers
        A[90, 40] = NAR
        A[120, 3] = NAR
        A[100, 7] = NAR
    for alpha in (ONE, 0x3A000000):
        want = B.copy(order="F")
        assert oracle.rtrsm("L", "L", "N", diag, m, n, alpha, A, want) == 0
        got = B.copy(order="F")
        pga.Rtrsm("L", "L", "N", diag, m, n, alpha, A, m, got, m)
        _same(got, want)


@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("special", [False, True])
def test_trsm_left_upper_tall(pga, oracle, diag, special):
    """Left upper solves with many rows (64-row diagonal blocks from the bottom up, then a
    descending-k update of the rows above each block)."""
    m, n = 200, 7
    A = _tri(oracle, 76, m, True)
    B = posit_matrix(oracle, 77, m, n)
    B[150, :] = 0  # zero b_k skip their step
    B[70, ::2] = 0
    if special:  # NaR in U under skipped and unskipped multipliers
        A[40, 150] = NAR
        A[10, 190] = NAR
        A[3, 64] = NAR
    for alpha in (ONE, 0x3A000000):
        want = B.copy(order="F")
        assert oracle.rtrsm("L", "U", "N", diag, m, n, alpha, A, want) == 0
        got = B.copy(order="F")
        pga.Rtrsm("L", "U", "N", diag, m, n, alpha, A, m, got, m)
        _same(got, want)


@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_left_lower_trans_chain(pga, oracle, diag):
    """Left, lower, transposed: one chain of additions per right-hand side, longer than
    the kernel's LDS chunk (4096 terms), with zero products, a NaR entry and sums beyond
    the decoded fast range (a huge right-hand-side entry)."""
    m, n = 4200, 4
    A = _tri(oracle, 74, m, False)
    B = posit_matrix(oracle, 75, m, n)
    A[4150:4190, 17] = 0  # zero products inside row 17's chain
    A[m - 5, 100] = NAR   # rows 100 and below-index rows of every column become NaR
    B[m - 3, 2] = oracle.from_double(1e35)  # column 2: sums beyond the fast range
    B[9, 3] = 0
    for alpha in (ONE, 0x3A000000):
        want = B.copy(order="F")
        assert oracle.rtrsm("L", "L", "T", diag, m, n, alpha, A, want) == 0
        got = B.copy(order="F")
        pga.Rtrsm("L", "L", "T", diag, m, n, alpha, A, m, got, m)
        _same(got, want)


def test_trsm_left_lower_trans_chain_long(pga, oracle):
    """Rows longer than the pipelined kernel's buffer (8192 terms): the solve that forms
    each row's terms in chunks of 4096 before wave 0 adds them."""
    m, n = 8300, 1
    A = _tri(oracle, 76, m, False)
    B = posit_matrix(oracle, 77, m, n)
    want = B.copy(order="F")
    assert oracle.rtrsm("L", "L", "T", "N", m, n, ONE, A, want) == 0
    got = B.copy(order="F")
    pga.Rtrsm("L", "L", "T", "N", m, n, ONE, A, m, got, m)
    _same(got, want)


@pytest.mark.parametrize("kind", ["short", "wide", "tiny", "cancel"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_left_lower_trans_chain_classes(pga, oracle, kind, diag):
    """The chain's integer grid steps on data that stresses them: short fractions (terms
    that tie between two posits of the running sum's binade, and exact sums), magnitudes
    spread over 2^-40 .. 2^40 (sums that leave their binade often), values near the ends
    of the posit range (scales outside -108 .. 107, summed exactly), and sums that cancel
    to zero."""
    m, n = 700, 3
    rng = np.random.default_rng({"short": 1, "wide": 2, "tiny": 3, "cancel": 4}[kind])
    if kind == "short":
        Ad = rng.integers(-8, 9, (m, m)) / 8.0
        Bd = rng.integers(-16, 17, (m, n)).astype(float)
    elif kind == "wide":
        Ad = rng.uniform(-1, 1, (m, m)) * 2.0 ** rng.integers(-40, 41, (m, m))
        Bd = rng.uniform(-1, 1, (m, n)) * 2.0 ** rng.integers(-20, 21, (m, n))
    elif kind == "tiny":
        Ad = rng.uniform(-1, 1, (m, m)) * 2.0 ** -100
        Bd = rng.uniform(-1, 1, (m, n)) * 2.0 ** rng.choice([-20, -8, 0], (m, n))
    else:
        Ad = np.ones((m, m))
        Bd = np.where(np.arange(m)[:, None] % 2 == 0, 1.0, -1.0) * np.ones((m, n))
    Ad[np.diag_indices(m)] = 1.0 + np.arange(m) % 3
    A = np.asfortranarray(oracle.from_double_v(Ad.ravel()).reshape(m, m))
    B = np.asfortranarray(oracle.from_double_v(Bd.ravel()).reshape(m, n))
    want = B.copy(order="F")
    assert oracle.rtrsm("L", "L", "T", diag, m, n, ONE, A, want) == 0
    got = B.copy(order="F")
    pga.Rtrsm("L", "L", "T", diag, m, n, ONE, A, m, got, m)
    _same(got, want)


@pytest.mark.parametrize("form", [("L", "N", "U"), ("L", "N", "N"), ("U", "N", "N"), ("U", "T", "N")])
@pytest.mark.parametrize("kind", ["plain", "special", "range"])
def test_trsm_left_columns_fp64(pga, oracle, form, kind):
    """Left solves with alpha = 1, more rows than a diagonal block and few right-hand
    sides: one workgroup per column in FP64 registers (1024 rows per register slot, 64
    rows of the sweep per barrier), a column redone at the pattern level when a value
    leaves the FP64 range.  1100 rows: two slots and a last block of 12 rows; zero
    entries of B (skipped steps), NaR in the triangle and in B, values near the ends of
    the posit range."""
    uplo, trans, diag = form
    m, n = 1100, 5
    A = _tri(oracle, 90, m, uplo == "U")
    B = posit_matrix(oracle, 91, m, n)
    B[40, :] = 0
    B[1090, ::2] = 0
    B[:, 4] = 0
    B[700, 4] = posit_matrix(oracle, 92, 1, 1)[0, 0]
    if kind == "special":
        r, c = (900, 130) if uplo == "L" else (130, 900)
        A[r, c] = NAR
        B[555, 1] = NAR
    elif kind == "range":
        B[1000, 2] = oracle.from_double(2.0 ** 100)
        B[10, 3] = oracle.from_double(2.0 ** -100)
    want = B.copy(order="F")
    assert oracle.rtrsm("L", uplo, trans, diag, m, n, ONE, A, want) == 0
    got = B.copy(order="F")
    pga.Rtrsm("L", uplo, trans, diag, m, n, ONE, A, m, got, m)
    _same(got, want)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_left_wide(pga, oracle, uplo, diag):
    """Left no-transpose solves with many right-hand sides (several columns per workgroup)."""
    m, n = 40, 2304
    A = _tri(oracle, 70, m, uplo == "U")
    B = posit_matrix(oracle, 71, m, n)
    B[5, ::7] = 0  # zero b_k entries skip their step
    B[30, 1::5] = 0
    for alpha in (ONE, 0x3A000000):
        want = B.copy(order="F")
        assert oracle.rtrsm("L", uplo, "N", diag, m, n, alpha, A, want) == 0
        got = B.copy(order="F")
        pga.Rtrsm("L", uplo, "N", diag, m, n, alpha, A, m, got, m)
        _same(got, want)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("alpha,beta", [(MONE, ONE), (0x3C000000, 0), (0, 0x44000000)])
def test_syrk(pga, oracle, uplo, trans, alpha, beta):
    n, k = 45, 70
    A = posit_matrix(oracle, 7, n if trans == "N" else k, k if trans == "N" else n)
    A[:, 5] = 0  # zero multipliers are skipped in the 'N' form
    C = posit_matrix(oracle, 8, n, n)
    want = C.copy(order="F")
    assert oracle.rsyrk(uplo, trans, n, k, alpha, A, beta, want) == 0
    got = C.copy(order="F")
    pga.Rsyrk(uplo, trans, n, k, alpha, A, A.shape[0], beta, got, n)
    _same(got, want)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_gemv(pga, oracle, trans, incx, incy):
    m, n = 50, 33
    A = posit_matrix(oracle, 9, m, n)
    lenx, leny = (n, m) if trans == "N" else (m, n)
    x = posit_matrix(oracle, 10, (lenx - 1) * abs(incx) + 1, 1)[:, 0].copy()
    y = posit_matrix(oracle, 11, (leny - 1) * abs(incy) + 1, 1)[:, 0].copy()
    for alpha, beta in [(ONE, 0), (MONE, ONE), (0x3A000000, 0xC4000000)]:
        want = y.copy()
        assert oracle.rgemv(trans, m, n, alpha, A, x, incx, beta, want, incy) == 0
        got = y.copy()
        pga.Rgemv(trans, m, n, alpha, A, m, x, incx, beta, got, incy)
        _same(got, want)


def test_scal_iamax_laswp(pga, oracle):
    rng = np.random.default_rng(3)
    x = posit_matrix(oracle, 12, 1000, 1)[:, 0].copy()
    want = x.copy()
    oracle.L.orc_rscal(1000 // 3, 0x3A000000, want.ctypes.data, 3)
    got = x.copy()
    pga.Rscal(1000 // 3, 0x3A000000, got, 3)
    _same(got, want)
    for trial in range(20):
        v = x.copy()
        if trial % 3 == 0:
            v[rng.integers(0, v.size, 4)] = NAR
        if trial % 3 == 1:  # ties: the first maximum wins
            v[[100, 500, 900]] = v[np.argmax(np.abs(oracle.to_double_v(v)))]
        inc = 1 + trial % 2
        n = v.size // inc
        assert pga.iRamax(n, v, inc) == oracle.iramax(v, inc, n)
    A = posit_matrix(oracle, 13, 60, 17)
    ipiv = rng.integers(1, 61, 20).astype(np.int32)
    ipiv = np.maximum(ipiv, np.arange(1, 21, dtype=np.int32))
    for incx in (1, -1):
        want = A.copy(order="F")
        oracle.L.orc_rlaswp(17, want.ctypes.data, 60, 3, 18, ipiv.ctypes.data, incx)
        got = A.copy(order="F")
        pga.Rlaswp(17, got, 60, 3, 18, ipiv, incx)
        _same(got, want)


@pytest.mark.parametrize("m,n", [(150, 150), (200, 120), (90, 160)])
@pytest.mark.parametrize("nb", [0, 1, 16, 64])
def test_getrf(pga, oracle, m, n, nb):
    A = posit_matrix(oracle, 20 + m + n, m, n)
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
    got = A.copy(order="F")
    ipiv = np.zeros(min(m, n), np.int32)
    info = pga.Rgetrf_opencl(m, n, got, m, ipiv, nb)
    assert info == winfo
    _same(ipiv, wpiv)
    _same(got, want)


@pytest.mark.parametrize("kind", ["zero_column", "nar", "rank_deficient"])
def test_getrf_special(pga, oracle, kind):
    n = 96
    A = posit_matrix(oracle, 30, n, n)
    if kind == "zero_column":
        A[:, 40] = 0
    elif kind == "nar":
        A[7, 3] = NAR
        A[50, 60] = NAR
    else:
        A[:, 10] = A[:, 20]
    for nb in (0, 8, 32):
        want = A.copy(order="F")
        _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
        got = A.copy(order="F")
        ipiv = np.zeros(n, np.int32)
        info = pga.Rgetrf_opencl(n, n, got, n, ipiv, nb)
        assert info == winfo, (nb, info, winfo)
        _same(ipiv, wpiv)
        _same(got, want)


def test_getrf2_and_getrs(pga, oracle):
    n, nrhs = 130, 5
    A = posit_matrix(oracle, 40, n, n)
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf2(want, nthreads=THREADS)
    got = A.copy(order="F")
    ipiv = np.zeros(n, np.int32)
    assert pga.Rgetrf2(n, n, got, n, ipiv) == winfo
    _same(ipiv, wpiv)
    _same(got, want)
    B = posit_matrix(oracle, 41, n, nrhs)
    for trans in ("N", "T"):
        wb = B.copy(order="F")
        assert oracle.rgetrs(trans, want, wpiv, wb) == (0, 0)
        gb = B.copy(order="F")
        assert pga.Rgetrs(trans, n, nrhs, got, n, ipiv, gb, n) == 0
        _same(gb, wb)


def _spd(oracle, seed, n):
    M = uniform_doubles(seed, n * n).reshape(n, n)
    S = M @ M.T / n + np.eye(n)
    return np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("nb", [0, 1, 16, 40])
def test_potrf_potrs(pga, oracle, uplo, nb):
    n, nrhs = 150, 3
    A = _spd(oracle, 50, n)
    want = A.copy(order="F")
    assert oracle.rpotrf(uplo, want, nb=nb, nthreads=THREADS) == (0, 0)
    got = A.copy(order="F")
    assert pga.Rpotrf_opencl(uplo, n, got, n, nb) == 0
    _same(got, want)
    B = posit_matrix(oracle, 51, n, nrhs)
    wb = B.copy(order="F")
    assert oracle.rpotrs(uplo, want, wb) == (0, 0)
    gb = B.copy(order="F")
    assert pga.Rpotrs(uplo, n, nrhs, got, n, gb, n) == 0
    _same(gb, wb)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_not_pd(pga, oracle, uplo):
    n = 100
    A = _spd(oracle, 52, n)
    A[61, 61] = oracle.from_double(-2.0)
    for nb in (0, 1, 16):
        want = A.copy(order="F")
        _, winfo = oracle.rpotrf(uplo, want, nb=nb, nthreads=THREADS)
        got = A.copy(order="F")
        info = pga.Rpotrf_opencl(uplo, n, got, n, nb)
        assert info == winfo and info > 0
        _same(got, want)  # the factorization stops at the same place


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("kind", ["huge", "tiny", "sparse", "nar"])
def test_potrf_diagonal_block_ranges(pga, oracle, kind, uplo):
    """The diagonal blocks (and the upper factor's row-block solves) run in FP64 while
    every product, quotient and sum stays in the fast scale range and fall back to the
    pattern-level form otherwise: scaled far up or down, with exact zeros (skipped
    multipliers in the lower form), or with a NaR (a failing block)."""
    n = 150
    if kind == "sparse":
        rng = np.random.default_rng(57)
        S = rng.uniform(-1, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.1)
        S = np.tril(S, -1)
        S = S + S.T + np.eye(n) * (n / 5 + 2)
    else:
        M = uniform_doubles(58, n * n).reshape(n, n)
        S = M @ M.T / n + np.eye(n)
        S *= {"huge": 2.0 ** 112, "tiny": 2.0 ** -116, "nar": 1.0}[kind]
    A = np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))
    if kind == "nar":
        A[100, 70] = A[70, 100] = 0x80000000
    for nb in (0, 16, 64):
        want = A.copy(order="F")
        _, winfo = oracle.rpotrf(uplo, want, nb=nb, nthreads=THREADS)
        got = A.copy(order="F")
        info = pga.Rpotrf_opencl(uplo, n, got, n, nb)
        assert info == winfo and (info > 0) == (kind == "nar")
        _same(got, want)


def test_potrf2(pga, oracle):
    n = 77
    A = _spd(oracle, 53, n)
    for uplo in "LU":
        want = A.copy(order="F")
        assert oracle.rpotrf2(uplo, want) == (0, 0)
        got = A.copy(order="F")
        assert pga.Rpotrf2(uplo, n, got, n) == 0
        _same(got, want)


def test_lu_1024_nb256_device(pga, oracle):
    """Device-resident blocked LU (config-3 blocking) at n = 1024, bit-exact, plus the
    solve with b = A*ones built by the posit GEMV."""
    n, nb = 1024, 256
    A = posit_matrix(oracle, 60, n, n)
    dA = pga.DeviceArray.from_host(A)
    dp = pga.DeviceArray(n, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(n, n, dA, n, dp, di, nb)
    pga.synchronize()
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
    assert int(di.to_host()[0]) == winfo
    _same(dp.to_host(), wpiv)
    _same(dA.to_host(), want)
    ones = np.full(n, ONE, np.uint32)
    b = np.zeros(n, np.uint32)
    oracle.rgemv("N", n, n, ONE, A, ones, 1, 0, b, 1)
    gb = b.copy()
    pga.Rgemv("N", n, n, ONE, A, n, ones, 1, 0, gb, 1)
    _same(gb, b)
    B = b.reshape(n, 1).copy(order="F")  # the oracle solves in place
    oracle.rgetrs("N", want, wpiv, B)
    dB = pga.DeviceArray.from_host(b.reshape(n, 1).copy(order="F"))
    pga.getrs_device("N", n, 1, dA, n, dp, dB, n)
    pga.synchronize()
    _same(dB.to_host(), B)


@pytest.mark.parametrize("m,n,nb,kind", [(1100, 1100, 64, None), (1300, 1100, 128, None), (1100, 1300, 96, None),
                                         (1100, 1100, 64, "zero_column"), (1040, 1040, 256, "nar")])
def test_lu_lookahead_device(pga, oracle, m, n, nb, kind):
    """From min(m, n) = 1024 the LU factors panel j+1 while a second stream updates the
    columns beyond it; pivots, info and every element must stay the same bits, for tall
    and wide shapes, a zero column and NaR entries."""
    A = posit_matrix(oracle, 64 + m + n, m, n)
    if kind == "zero_column":
        A[:, 700] = 0
    elif kind == "nar":
        A[5, 3] = NAR
        A[900, 600] = NAR
    mn = min(m, n)
    dA = pga.DeviceArray.from_host(A)
    dp = pga.DeviceArray(mn, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(m, n, dA, m, dp, di, nb)
    pga.synchronize()
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
    assert int(di.to_host()[0]) == winfo
    if kind == "zero_column":
        assert winfo == 701
    _same(dp.to_host(), wpiv)
    _same(dA.to_host(), want)


@pytest.mark.parametrize("m,n,nb,kind", [(2600, 520, 128, None), (4100, 300, 256, "zero_column"),
                                         (3000, 200, 64, "nar"), (2100, 130, 64, "ties"),
                                         (2200, 90, 32, "nar_pivot"), (9000, 260, 256, None)])
def test_lu_pair_multi_workgroup(pga, oracle, m, n, nb, kind):
    """Two-column nodes with at least 2048 rows run on several workgroups that meet at a
    grid barrier per pivot search (2 to 16 blocks of rows): pivots found in any block,
    ties between blocks (the first row wins), zero columns (no interchange of the pivot
    column, info), NaR entries, a NaR on the diagonal over zeros (pivot below it with value
    zero), and pending node solves after a pair."""
    A = posit_matrix(oracle, 91 + m + n, m, n)
    if kind == "zero_column":
        A[:, 130] = 0
    elif kind == "nar":
        A[5, 3] = NAR
        A[2900, 100] = NAR
    elif kind == "ties":
        A[:, 40] = oracle.from_double(0.5)  # equal magnitudes in every block
        A[1500:, 41] = oracle.from_double(-0.75)
    elif kind == "nar_pivot":
        A[:, 20] = 0
        A[20, 20] = NAR
    mn = min(m, n)
    dA = pga.DeviceArray.from_host(A)
    dp = pga.DeviceArray(mn, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(m, n, dA, m, dp, di, nb)
    pga.synchronize()
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
    assert int(di.to_host()[0]) == winfo
    _same(dp.to_host(), wpiv)
    _same(dA.to_host(), want)


@pytest.mark.parametrize("m,n,nb,kind", [(1100, 200, 64, "huge"), (1100, 200, 64, "tiny"),
                                         (3000, 128, 64, "nar_late"), (1500, 96, 96, "zero_u_nar"),
                                         (600, 64, 64, "ties"), (8400, 64, 64, None), (300, 40, 40, None)])
def test_lu_panel_nodes(pga, oracle, m, n, nb, kind):
    """32-column recursion nodes as one multi-workgroup launch (1 to 32 workgroups, FP64
    multiply-adds in registers): scales beyond 2^+-108 in the middle of a node (the node
    starts again with the pattern-level operations), a NaR that first meets the FP64 form
    at the 20th column of a node, a zero multiplier above a NaR (the recursion's Rtrsm
    rows skip it, its Rgemm rows do not), ties between workgroups, two workgroups per lane
    slot (8400 rows), and a node width that is not a power of two (40 columns: the usual
    recursion)."""
    A = posit_matrix(oracle, 7 + m + n, m, n)
    if kind == "huge":
        A[700:720, 40:60] = oracle.from_double(2.0 ** 100)
    elif kind == "tiny":
        A[900:905, 70] = oracle.from_double(2.0 ** -100)
    elif kind == "nar_late":
        A[2500, 84] = NAR  # column 20 of the third node
    elif kind == "zero_u_nar":
        A[:, 2] = 0
        A[2, 2] = oracle.from_double(1.0)
        A[3:, 2] = 0
        A[1400, 1] = NAR  # below the diagonal of column 1: in L
        A[1, 5:30] = 0    # row 1 is in U: zero multipliers for columns 5 .. 29
    elif kind == "ties":
        A[:, 10] = oracle.from_double(-0.5)
    mn = min(m, n)
    dA = pga.DeviceArray.from_host(A)
    dp = pga.DeviceArray(mn, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(m, n, dA, m, dp, di, nb)
    pga.synchronize()
    want = A.copy(order="F")
    _, wpiv, winfo = oracle.rgetrf(want, nb=nb, nthreads=THREADS)
    assert int(di.to_host()[0]) == winfo
    _same(dp.to_host(), wpiv)
    _same(dA.to_host(), want)


def test_cholesky_1024_device(pga, oracle):
    n = 1024
    A = _spd(oracle, 61, n)
    dA = pga.DeviceArray.from_host(A)
    di = pga.DeviceArray(1, np.int32, "C")
    pga.potrf_device("L", n, dA, n, di, 0)
    pga.synchronize()
    want = A.copy(order="F")
    assert oracle.rpotrf("L", want, nb=0, nthreads=THREADS) == (0, 0)
    assert int(di.to_host()[0]) == 0
    _same(np.tril(dA.to_host()), np.tril(want))


@pytest.mark.parametrize("uplo,n,nb,bad", [("L", 1100, 64, 0), ("L", 1100, 64, 700), ("U", 1100, 64, 0),
                                           ("U", 1100, 64, 700), ("L", 1300, 128, 0), ("L", 1300, 128, 5),
                                           ("L", 1300, 128, 1290), ("L", 1111, 48, 1063), ("L", 1040, 80, 0),
                                           ("L", 1040, 80, 959), ("L", 1030, 50, 0), ("L", 1030, 50, 777),
                                           ("U", 1300, 128, 5), ("U", 1300, 64, 1290), ("U", 1111, 48, 1063),
                                           ("U", 1040, 80, 0), ("U", 1030, 50, 777), ("U", 1024, 64, 63)])
def test_cholesky_panel_overlap(pga, oracle, uplo, n, nb, bad):
    """Orders from 1024 run on two streams, right-looking with one block of lookahead: the
    lower factor updates the trailing matrix in place, the upper one keeps each later
    element's Rsyrk / RgemmG sum in a workspace until its block row comes up.  The factor,
    and after a failing block the whole matrix the reference leaves behind, must stay the
    same bits."""
    A = _spd(oracle, 62, n)
    if bad:
        A[bad, bad] = oracle.from_double(-2.0)
    want = A.copy(order="F")
    _, winfo = oracle.rpotrf(uplo, want, nb=nb, nthreads=THREADS)
    dA = pga.DeviceArray.from_host(A)
    di = pga.DeviceArray(1, np.int32, "C")
    pga.potrf_device(uplo, n, dA, n, di, nb)
    pga.synchronize()
    assert int(di.to_host()[0]) == winfo == (bad + 1 if bad else 0)
    _same(dA.to_host(), want)


def test_full_size_invariants(pga):
    """The n = 8192 configurations through properties that do not need the oracle: the LU
    with nb = 256 equals the LU with nb = 64 bit for bit (pivots included), the lower
    Cholesky with nb = 64 equals the one with nb = 128, and both solves have small FP64
    backward errors."""
    n = 8192
    ONE_ = 0x40000000
    dA0 = pga.DeviceArray((n, n))
    pga.fill_uniform(dA0, 4242)
    ones = pga.DeviceArray.from_host(np.full(n, ONE_, np.uint32))
    db = pga.DeviceArray(n, np.uint32, "C")
    pga.gemv_device("N", n, n, ONE_, dA0, n, ones, 1, 0, db, 1)
    dp = pga.DeviceArray(n, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    dA = pga.DeviceArray((n, n))
    dA.copy_from(dA0)
    pga.getrf_device(n, n, dA, n, dp, di, 256)
    pga.synchronize()
    F256, P256 = dA.to_host(), dp.to_host()
    assert int(di.to_host()[0]) == 0
    dx = pga.DeviceArray((n, 1))
    dx.copy_from(db)
    pga.getrs_device("N", n, 1, dA, n, dp, dx, n)
    assert pga.backward_error_device(n, dA0, n, dx, db) < 1e-4
    dA.copy_from(dA0)
    pga.getrf_device(n, n, dA, n, dp, di, 64)
    pga.synchronize()
    _same(dp.to_host(), P256)
    _same(dA.to_host(), F256)
    del F256
    # SPD input on the device: S = M M^T / n + I is built in FP64 on the host
    M = uniform_doubles(4300, n * n).reshape(n, n)
    S = M @ M.T
    S /= n
    S[np.diag_indices(n)] += 1.0
    del M
    dS = pga.DeviceArray.from_host(np.asfortranarray(S))
    del S
    dC0 = pga.DeviceArray((n, n))
    pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, dC0.ptr, n * n, None), "from_f64")
    dS.free()
    pga.gemv_device("N", n, n, ONE_, dC0, n, ones, 1, 0, db, 1)
    dC = pga.DeviceArray((n, n))
    dC.copy_from(dC0)
    pga.potrf_device("L", n, dC, n, di, 64)
    pga.synchronize()
    assert int(di.to_host()[0]) == 0
    L64 = np.tril(dC.to_host())
    dx.copy_from(db)
    pga.potrs_device("L", n, 1, dC, n, dx, n)
    assert pga.backward_error_device(n, dC0, n, dx, db) < 1e-6
    dC.copy_from(dC0)
    pga.potrf_device("L", n, dC, n, di, 128)
    pga.synchronize()
    _same(np.tril(dC.to_host()), L64)
