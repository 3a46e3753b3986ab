"""Pins for the oracle's BLAS / LAPACK restatements (CPU only).

Each check is independent of the oracle's own code: brute force with exact rational
arithmetic (posit_exact) on tiny inputs, algebraic identities between the four GEMM
variants, invariances of the blocked factorizations, FP64 residuals, and constructed
info cases.
"""
from fractions import Fraction

import numpy as np
import pytest

import posit_exact as X
from datagen import posit_matrix

ONE, MONE, ZERO, NAR = 0x40000000, 0xC0000000, 0, 0x80000000


def _neg(p):
    return (-p) & 0xFFFFFFFF


def _brute_gemm(ta, tb, alpha, A, B, beta, C):
    """Rgemm's element definition (Rgemm.cpp:163-246) with exact-rational posit ops."""
    m, n = C.shape
    k = A.shape[0] if ta else A.shape[1]
    out = C.copy()
    for j in range(n):
        for i in range(m):
            def a(l):
                return int(A[l, i] if ta else A[i, l])

            def b(l):
                return int(B[j, l] if tb else B[l, j])

            if not ta:
                c = 0 if beta == 0 else (int(C[i, j]) if beta == ONE else X.mul(beta, int(C[i, j])))
                for l in range(k):
                    c = X.add(c, X.mul(X.mul(alpha, b(l)), a(l)))
            else:
                t = 0
                for l in range(k):
                    t = X.add(t, X.mul(a(l), b(l)))
                c = X.mul(alpha, t) if beta == 0 else X.add(X.mul(alpha, t), X.mul(beta, int(C[i, j])))
            out[i, j] = c
    return out


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("alpha,beta", [(ONE, 0), (MONE, ONE), (0x3A000000, 0xC8000000)])
def test_gemm_brute_force(oracle, ta, tb, alpha, beta):
    m, n, k = 5, 4, 6
    A = posit_matrix(oracle, 1, k if ta else m, m if ta else k, -2, 2)
    B = posit_matrix(oracle, 2, n if tb else k, k if tb else n, -2, 2)
    C = posit_matrix(oracle, 3, m, n, -2, 2)
    want = _brute_gemm(ta, tb, alpha, A, B, beta, C)
    got = C.copy(order="F")
    assert oracle.rgemm("T" if ta else "N", "T" if tb else "N", m, n, k, alpha, A, B, beta, got) == 0
    assert np.array_equal(got, want)


def test_gemm_variant_identities(oracle):
    """With alpha=1, beta=0 all four variants sum the same products in the same order."""
    m, n, k = 37, 29, 41
    A = posit_matrix(oracle, 10, m, k)
    B = posit_matrix(oracle, 11, k, n)
    res = {}
    for ta in "NT":
        for tb in "NT":
            a = np.asfortranarray(A.T) if ta == "T" else A
            b = np.asfortranarray(B.T) if tb == "T" else B
            C = np.zeros((m, n), np.uint32, order="F")
            assert oracle.rgemm(ta, tb, m, n, k, ONE, a, b, 0, C) == 0
            res[ta + tb] = C
    for v in ("NT", "TN", "TT"):
        assert np.array_equal(res["NN"], res[v]), v


def test_gemm_ascending_order(oracle):
    """Products [1, 2^30, -2^30]: ascending order gives (1 + 2^30) - 2^30 = 0, any order
    that adds 1 last gives 1."""
    p = [X.encode(Fraction(1)), X.encode(Fraction(2) ** 30), X.encode(-Fraction(2) ** 30)]
    A = np.asfortranarray(np.array([p], np.uint32))  # 1 x 3
    B = np.asfortranarray(np.full((3, 1), ONE, np.uint32))
    C = np.zeros((1, 1), np.uint32, order="F")
    oracle.rgemm("N", "N", 1, 1, 3, ONE, A, B, 0, C)
    assert C[0, 0] == 0
    A2 = np.asfortranarray(np.array([[p[1], p[2], p[0]]], np.uint32))
    C[:] = 0
    oracle.rgemm("N", "N", 1, 1, 3, ONE, A2, B, 0, C)
    assert C[0, 0] == ONE


def test_gemm_identity_and_special(oracle):
    n = 16
    I = np.zeros((n, n), np.uint32, order="F")
    np.fill_diagonal(I, ONE)
    B = posit_matrix(oracle, 5, n, n)
    C = np.full((n, n), NAR, np.uint32, order="F")  # beta == 0 never reads C
    oracle.rgemm("N", "N", n, n, n, ONE, I, B, 0, C)
    assert np.array_equal(C, B)
    # alpha == 0: C = beta*C (Rgemm.cpp:142-157)
    C = B.copy(order="F")
    oracle.rgemm("N", "N", n, n, n, 0, I, B, MONE, C)
    assert np.array_equal(C, (-B.astype(np.int64) & 0xFFFFFFFF).astype(np.uint32))
    # quick return: k == 0 and beta == 1 leaves C untouched
    C = np.full((n, n), NAR, np.uint32, order="F")
    oracle.rgemm("N", "N", n, n, 0, ONE, I, B, ONE, C)
    assert np.all(C == NAR)


def test_gemm_argument_errors(oracle):
    A = np.zeros((4, 4), np.uint32, order="F")
    assert oracle.rgemm("X", "N", 4, 4, 4, ONE, A, A, 0, A.copy(order="F")) == 1
    assert oracle.rgemm("N", "q", 4, 4, 4, ONE, A, A, 0, A.copy(order="F")) == 2
    assert oracle.rgemm("N", "N", -1, 4, 4, ONE, A, A, 0, A.copy(order="F")) == 3
    assert oracle.rgemm("N", "N", 4, 4, 4, ONE, A, A, 0, A.copy(order="F"), lda=3) == 8
    assert oracle.rgemm("N", "N", 4, 4, 4, ONE, A, A, 0, A.copy(order="F"), ldb=2) == 10
    assert oracle.rgemm("N", "N", 4, 4, 4, ONE, A, A, 0, A.copy(order="F"), ldc=1) == 13


def test_gemm_close_to_fp64(oracle):
    m, n, k = 64, 48, 256
    A = posit_matrix(oracle, 20, m, k)
    B = posit_matrix(oracle, 21, k, n)
    C = posit_matrix(oracle, 22, m, n)
    ref = oracle.to_double_v(C) - oracle.to_double_v(A) @ oracle.to_double_v(B)
    oracle.rgemm("N", "N", m, n, k, MONE, A, B, ONE, C)
    err = np.abs(oracle.to_double_v(C) - ref).max()
    assert err < 1e-6, err


# ---------------------------------------------------------------------------------
# LU
# ---------------------------------------------------------------------------------
def _brute_lu(A):
    """Textbook right-looking LU with partial pivoting, posit ops in the reference's
    order: pivot = first max of |a| (signed pattern order), l = (1/p)*a, then
    a_ic = a_ic + (-u_jc)*l_ij for ascending j."""
    A = [[int(v) for v in row] for row in A]
    n = len(A)
    ipiv, info = [], 0
    for j in range(n):
        col = [X.decode(0) for _ in range(n)]
        best, bi = None, j
        for i in range(j, n):
            v = A[i][j]
            av = (-v) & 0xFFFFFFFF if (v >> 31) else v
            sv = av - (1 << 32) if av >> 31 else av
            if best is None or sv > best:
                best, bi = sv, i
        ipiv.append(bi + 1)
        if A[bi][j] == 0:
            if info == 0:
                info = j + 1
        else:
            A[j], A[bi] = A[bi], A[j]
            r = X.div(ONE, A[j][j])
            for i in range(j + 1, n):
                A[i][j] = X.mul(r, A[i][j])
        for i in range(j + 1, n):
            for c in range(j + 1, n):
                A[i][c] = X.add(A[i][c], X.mul(_neg(A[j][c]), A[i][j]))
    return np.array(A, np.uint32), np.array(ipiv, np.int32), info


@pytest.mark.parametrize("nb", [0, 1, 2, 3])
def test_lu_brute_force(oracle, nb):
    n = 7
    A = posit_matrix(oracle, 31, n, n)
    want, wpiv, winfo = _brute_lu(A)
    got = A.copy(order="F")
    rc, ipiv, info = oracle.rgetrf(got, nb=nb)
    assert rc == 0 and info == winfo
    assert np.array_equal(ipiv, wpiv)
    assert np.array_equal(got, want)


def test_lu_block_size_invariance(oracle):
    n = 150
    A = posit_matrix(oracle, 40, n, n)
    ref = A.copy(order="F")
    _, piv0, info0 = oracle.rgetrf2(ref)
    for nb in (2, 3, 8, 16, 32, 64, 100):
        B = A.copy(order="F")
        _, piv, info = oracle.rgetrf(B, nb=nb)
        assert info == info0
        assert np.array_equal(piv, piv0), nb
        assert np.array_equal(B, ref), nb


def _plu_residual(oracle, A, F, ipiv):
    n = A.shape[0]
    LU = oracle.to_double_v(F)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = oracle.to_double_v(A).copy()
    for i, p in enumerate(ipiv):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    return np.abs(L @ U - PA).max() / np.abs(PA).max()


def test_lu_residual_and_solve(oracle):
    n = 200
    A = posit_matrix(oracle, 41, n, n)
    F = A.copy(order="F")
    _, ipiv, info = oracle.rgetrf(F, nb=32)
    assert info == 0
    assert _plu_residual(oracle, A, F, ipiv) < 1e-6
    # b = A * ones via posit GEMV (Rgemv N form), then solve
    x1 = np.full(n, ONE, np.uint32)
    b = np.zeros(n, np.uint32)
    assert oracle.rgemv("N", n, n, ONE, A, x1, 1, 0, b, 1) == 0
    B = np.asfortranarray(b.reshape(n, 1))
    rc, info = oracle.rgetrs("N", F, ipiv, B)
    assert rc == 0 and info == 0
    x = oracle.to_double_v(B[:, 0])
    assert np.abs(x - 1).max() < 1e-3


def test_lu_singular_info(oracle):
    n = 40
    A = posit_matrix(oracle, 42, n, n)
    A[:, 17] = 0  # column 18 (1-based) is zero: first zero pivot there
    F = A.copy(order="F")
    rc, ipiv, info = oracle.rgetrf(F, nb=8)
    assert rc == 0 and info == 18
    F2 = A.copy(order="F")
    _, _, info2 = oracle.rgetrf2(F2)
    assert info2 == 18 and np.array_equal(F, F2)


def test_lapack_argument_errors(oracle):
    A = np.zeros((4, 4), np.uint32, order="F")
    ipiv = np.zeros(4, np.int32)
    info = np.zeros(1, np.int32)
    import ctypes
    P = ctypes.c_void_p
    assert oracle.L.orc_rgetrf(-1, 4, P(A.ctypes.data), 4, P(ipiv.ctypes.data), P(info.ctypes.data), 0, 1) == -1
    assert oracle.L.orc_rgetrf(4, 4, P(A.ctypes.data), 2, P(ipiv.ctypes.data), P(info.ctypes.data), 0, 1) == -4
    assert oracle.L.orc_rpotrf(b"X", 4, P(A.ctypes.data), 4, P(info.ctypes.data), 0, 1) == -1


# ---------------------------------------------------------------------------------
# Cholesky
# ---------------------------------------------------------------------------------
def spd_matrix(oracle, seed, n):
    """M*M^T/n + I in FP64 from uniform M, converted to posit32."""
    from datagen import uniform_doubles

    M = uniform_doubles(seed, n * n).reshape(n, n)
    S = M @ M.T / n + np.eye(n)
    return np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))


def _brute_chol_lower(A):
    """Left-looking Cholesky, lower: d = a_jj + sum_k (-l_jk)*l_jk, l_jj = sqrt(d);
    l_ij = (1/l_jj) * (a_ij + sum_k (-l_jk)*l_ik), sums in ascending k."""
    n = A.shape[0]
    L = [[0] * n for _ in range(n)]
    for j in range(n):
        d = int(A[j, j])
        for k in range(j):
            if L[j][k] != 0:  # Rsyrk skips zero multipliers (Rsyrk.cpp:169)
                d = X.add(d, X.mul(_neg(L[j][k]), L[j][k]))
        if d == 0 or (d >> 31):
            return None, j + 1
        L[j][j] = X.sqrt(d)
        r = X.div(ONE, L[j][j])
        for i in range(j + 1, n):
            v = int(A[i, j])
            for k in range(j):
                v = X.add(v, X.mul(_neg(L[j][k]), L[i][k]))
            L[i][j] = X.mul(r, v)
    return L, 0


@pytest.mark.parametrize("nb", [0, 1, 2, 3])
def test_cholesky_brute_force(oracle, nb):
    n = 7
    A = spd_matrix(oracle, 50, n)
    L, info = _brute_chol_lower(A)
    assert info == 0
    F = A.copy(order="F")
    rc, info = oracle.rpotrf("L", F, nb=nb)
    assert rc == 0 and info == 0
    for i in range(n):
        for j in range(i + 1):
            assert int(F[i, j]) == L[i][j], (i, j)


def test_cholesky_lower_block_size_invariance(oracle):
    n = 130
    A = spd_matrix(oracle, 51, n)
    ref = A.copy(order="F")
    oracle.rpotrf2("L", ref)
    for nb in (2, 5, 16, 64, 100):
        F = A.copy(order="F")
        rc, info = oracle.rpotrf("L", F, nb=nb)
        assert info == 0
        assert np.array_equal(np.tril(F), np.tril(ref)), nb


def test_cholesky_residual_solve_and_upper(oracle):
    n = 160
    A = spd_matrix(oracle, 52, n)
    Ad = oracle.to_double_v(A)
    for uplo in "LU":
        F = A.copy(order="F")
        rc, info = oracle.rpotrf(uplo, F, nb=0)
        assert rc == 0 and info == 0
        T = oracle.to_double_v(F)
        T = np.tril(T) if uplo == "L" else np.triu(T)
        R = T @ T.T if uplo == "L" else T.T @ T
        assert np.abs(R - Ad).max() / np.abs(Ad).max() < 1e-6
        B = np.asfortranarray(oracle.from_double_v(Ad @ np.ones(n)).reshape(n, 1))
        rc, info = oracle.rpotrs(uplo, F, B)
        assert rc == 0 and info == 0
        assert np.abs(oracle.to_double_v(B[:, 0]) - 1).max() < 1e-3


def _unblocked_right_looking_lower(oracle, A):
    """Lower Cholesky one column at a time, right-looking: column j final (test, sqrt,
    scale by 1/l_jj), then its terms -l_kj*l_ij applied to every later column k, each
    skipped when the multiplier l_kj is zero.  The order of the device's register-resident
    diagonal block; None when a diagonal entry fails the test."""
    n = A.shape[0]
    A = A.copy(order="F")
    for j in range(n):
        v = int(A[j, j])
        if v == 0 or (v >> 31):
            return None
        d = oracle.sqrt(v)
        A[j, j] = d
        if j + 1 == n:
            break
        r = oracle.div(ONE, d)
        A[j + 1:, j] = oracle.mul_v(np.full(n - j - 1, r, dtype=np.uint32), A[j + 1:, j])
        for k in range(j + 1, n):
            m = int(A[k, j])
            if m == 0:
                continue
            A[k:, k] = oracle.sub_v(A[k:, k], oracle.mul_v(np.full(n - k, m, dtype=np.uint32), A[k:, j]))
    return A


@pytest.mark.parametrize("n,kind", [(64, "dense"), (37, "dense"), (64, "sparse"), (50, "sparse"), (1, "dense")])
def test_cholesky_lower_unblocked_order(oracle, n, kind):
    """Rpotrf2 'L' (recursive) gives each element the same terms in the same order as the
    column-at-a-time right-looking form, zero-multiplier skips included."""
    if kind == "dense":
        A = spd_matrix(oracle, 55, n)
    else:  # diagonally dominant with a sparse pattern: exact zeros in the factor
        rng = np.random.default_rng(56 + n)
        S = rng.uniform(-1, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.15)
        S = np.tril(S, -1)
        S = S + S.T + np.eye(n) * (n / 4 + 2)
        A = np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))
    want = A.copy(order="F")
    rc, info = oracle.rpotrf2("L", want)
    assert rc == 0 and info == 0
    got = _unblocked_right_looking_lower(oracle, A)
    assert got is not None
    assert np.array_equal(got, want)
    if kind == "sparse":
        assert (got[np.tril_indices(n, -1)] == 0).sum() > n  # the skips were exercised


def test_cholesky_not_pd_info(oracle):
    n = 70
    A = spd_matrix(oracle, 53, n)
    A[40, 40] = oracle.from_double(-1.0)  # leading minor 41 is not positive definite
    F = A.copy(order="F")
    rc, info = oracle.rpotrf("L", F, nb=16)
    assert rc == 0 and info == 41


def _right_looking_lower(oracle, A, nb):
    """Lower Cholesky in right-looking order, from the oracle's Rpotrf2 / Rtrsm / Rsyrk /
    Rgemm: once block j is factored, its columns update every later column at once
    (Rsyrk 'L','N' on each diagonal block, Rgemm 'N','T' below it, K = block j only);
    after a failing block the saved input is put back over the panel below it and every
    later column.  This is the order of the device's lower driver for large n."""
    n = A.shape[0]
    save = A.copy(order="F")
    info = 0
    for j in range(0, n, nb):
        jb = min(nb, n - j)
        D = np.asfortranarray(A[j:j + jb, j:j + jb])
        _, f = oracle.rpotrf2("L", D)
        A[j:j + jb, j:j + jb] = D
        if f:
            info = j + f
            break
        if j + jb == n:
            break
        P = np.asfortranarray(A[j + jb:, j:j + jb])
        oracle.rtrsm("R", "L", "T", "N", n - j - jb, jb, ONE, D, P)
        A[j + jb:, j:j + jb] = P
        for c in range(j + jb, n, nb):
            cb = min(nb, n - c)
            Lc = np.asfortranarray(A[c:c + cb, j:j + jb])
            C = np.asfortranarray(A[c:c + cb, c:c + cb])
            oracle.rsyrk("L", "N", cb, jb, MONE, Lc, ONE, C)
            A[c:c + cb, c:c + cb] = C
            if c + cb < n:
                Lr = np.asfortranarray(A[c + cb:, j:j + jb])
                C = np.asfortranarray(A[c + cb:, c:c + cb])
                oracle.rgemm("N", "T", n - c - cb, cb, jb, MONE, Lr, Lc, ONE, C)
                A[c + cb:, c:c + cb] = C
    if info:
        J = (info - 1) // nb * nb
        r0 = min(n, J + nb)
        for c in range(J, n):
            A[max(c, r0):, c] = save[max(c, r0):, c]
    return info


@pytest.mark.parametrize("n,nb,bad", [(150, 32, None), (150, 32, 0), (150, 32, 40), (150, 32, 100),
                                      (150, 32, 149), (130, 48, None), (130, 48, 97)])
def test_cholesky_lower_right_looking_order(oracle, n, nb, bad):
    """The lower factor splits by K block: each element gets the same terms in the same
    order whether its columns are applied left-looking (Rpotrf_opencl) or one finished
    block at a time, and after a failure the matrix the reference leaves is recovered from
    the saved input."""
    A = spd_matrix(oracle, 54, n)
    if bad is not None:
        A[bad, bad] = oracle.from_double(-2.0)
    F = A.copy(order="F")
    _, want = oracle.rpotrf("L", F, nb=nb)
    G = A.copy(order="F")
    got = _right_looking_lower(oracle, G, nb)
    assert got == want == (0 if bad is None else bad + 1)
    assert np.array_equal(G, F)
