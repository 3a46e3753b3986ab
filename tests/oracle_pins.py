"""Independent checks of the oracle's BLAS / LAPACK restatements, shared by
tests/test_oracle_pins.py (the oracle passes every check) and tests/test_oracle_mutants.py
(a copy of the oracle with one planted mistake fails at least one check).

Each check computes the expected result on tiny inputs with exact rational posit
arithmetic (posit_exact: decode to a rational, operate exactly, round the bit string) and
states each routine element by element: which terms an element receives, in which order,
which of them are skipped, and whether a diagonal enters as a quotient or as a reciprocal
times the value, as the reference's loops define it (file:line at each function).  The
oracle is the code under test; its results are only compared.

Every check returns a list of failure descriptions (empty when the oracle agrees)."""
import numpy as np

import posit_exact as X

ONE, MONE, ZERO, NAR = 0x40000000, 0xC0000000, 0, 0x80000000
ALPHAS = (ONE, MONE, 0x3A000000, 0)  # 1, -1, 0.625, 0


def neg(p):
    return (-p) & 0xFFFFFFFF


def sub(a, b):
    return X.add(a, neg(b))


def _pos(vals):
    return [X.from_double(float(v)) for v in vals]


def _matrix(rng, r, c, lo=-2.0, hi=2.0, zeros=0.0, nars=0.0):
    """r x c list of pattern rows, uniform in [lo, hi) with exact zeros / NaRs sprinkled."""
    M = [_pos(rng.uniform(lo, hi, c)) for _ in range(r)]
    for i in range(r):
        for j in range(c):
            u = rng.random()
            if u < zeros:
                M[i][j] = ZERO
            elif u < zeros + nars:
                M[i][j] = NAR
    return M


def _np(M):
    return np.asfortranarray(np.array(M, dtype=np.uint32).reshape(len(M), -1))


def _diff(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want, dtype=got.dtype).reshape(got.shape)
    bad = np.argwhere(got != want)
    if bad.size == 0:
        return []
    i = tuple(bad[0])
    return [f"{name}: {len(bad)} elements differ, first {i}: oracle {int(got[i]):#010x} exact {int(want[i]):#010x}"]


# ---------------------------------------------------------------------------------------
# Rtrsm (Rtrsm.cpp:31-291)
# ---------------------------------------------------------------------------------------
def trsm_exact(side, uplo, transa, diag, alpha, A, B):
    """Element-by-element definition of the reference's triangular solves."""
    m, n = len(B), len(B[0])
    nounit = diag == "N"
    if alpha == ZERO:  # :101-107
        return [[ZERO] * n for _ in range(m)]
    X_ = [row[:] for row in B]
    if side == "L" and transa == "N":
        # :120-156, per column: start from alpha*b (when alpha != 1); the unknowns are
        # settled in sweep order (upper: last row first, lower: first row first); unknown k
        # receives  - x_p * a_kp  from every settled p in sweep order whose value was
        # non-zero when it was settled (:128, :146 skip zeros), then, if non-zero and
        # non-unit, becomes value / a_kk (a quotient, :130, :148)
        order = range(m - 1, -1, -1) if uplo == "U" else range(m)
        for j in range(n):
            v = [X.mul(alpha, X_[i][j]) if alpha != ONE else X_[i][j] for i in range(m)]
            settled = []
            for k in order:
                t = v[k]
                for p in settled:
                    if v[p] != ZERO:
                        t = sub(t, X.mul(v[p], A[k][p]))
                if t != ZERO and nounit:
                    t = X.div(t, A[k][k])
                v[k] = t
                settled.append(k)
            for i in range(m):
                X_[i][j] = v[i]
        return X_
    if side == "L":
        # :161-186, per column: x_i = (alpha*b_i - a_ki*x_k ...) with the terms of the
        # settled unknowns k in ascending k (upper: k < i; lower: k > i, settled first),
        # no skips, then divided by a_ii when non-unit
        for j in range(n):
            x = [None] * m
            order = range(m) if uplo == "U" else range(m - 1, -1, -1)
            for i in order:
                t = X.mul(alpha, B[i][j])
                ks = range(i) if uplo == "U" else range(i + 1, m)
                for k in ks:
                    t = sub(t, X.mul(A[k][i], x[k]))
                if nounit:
                    t = X.div(t, A[i][i])
                x[i] = t
            for i in range(m):
                X_[i][j] = x[i]
        return X_
    if transa == "N":
        # :194-236, per row of B: columns settled in order (upper: ascending, lower:
        # descending); column j starts from alpha*b_ij (when alpha != 1), receives
        # - a_kj * x_ik from the settled columns k in ascending k, skipping a_kj == 0
        # (:202, :223), then is multiplied by the reciprocal 1/a_jj when non-unit (:206-209)
        for i in range(m):
            x = [None] * n
            cols = range(n) if uplo == "U" else range(n - 1, -1, -1)
            for j in cols:
                t = X.mul(alpha, B[i][j]) if alpha != ONE else B[i][j]
                ks = range(j) if uplo == "U" else range(j + 1, n)
                for k in ks:
                    if A[k][j] != ZERO:
                        t = sub(t, X.mul(A[k][j], x[k]))
                if nounit:
                    t = X.mul(X.div(ONE, A[j][j]), t)
                x[j] = t
            X_[i] = x
        return X_
    # :241-285, per row of B: columns settled in order (upper: descending, lower:
    # ascending); when column k settles, y_k = (1/a_kk) * its value (non-unit, a
    # reciprocal), every later column j then receives - a_jk * y_k unless a_jk == 0
    # (:250, :272), and only after that is the stored x_k = alpha * y_k (alpha != 1)
    for i in range(m):
        y = [None] * n
        t = list(B[i])
        cols = range(n - 1, -1, -1) if uplo == "U" else range(n)
        for k in cols:
            v = t[k]
            if nounit:
                v = X.mul(X.div(ONE, A[k][k]), v)
            y[k] = v
            others = range(k) if uplo == "U" else range(k + 1, n)
            for j in others:
                if A[j][k] != ZERO:
                    t[j] = sub(t[j], X.mul(A[j][k], v))
        X_[i] = [X.mul(alpha, y[j]) if alpha != ONE else y[j] for j in range(n)]
    return X_


def _tri_matrix(rng, n, uplo, nar_off=False):
    """A well-scaled triangular matrix (diagonal in [1, 2)); off-diagonal entries of the
    unused triangle are NaR so that reading them shows."""
    A = _matrix(rng, n, n, -1.0, 1.0, zeros=0.25)
    for i in range(n):
        A[i][i] = X.from_double(float(rng.uniform(1.0, 2.0)) * (1 if rng.random() < 0.7 else -1))
        for j in range(n):
            if (uplo == "U" and i > j) or (uplo == "L" and i < j):
                A[i][j] = NAR
    if nar_off:  # one NaR inside the used triangle: only skipped products keep it out
        i, j = (0, n - 1) if uplo == "U" else (n - 1, 0)
        A[i][j] = NAR
    return A


def check_trsm(oracle, seed=1):
    rng = np.random.default_rng(seed)
    fails = []
    for side in "LR":
        for uplo in "UL":
            for transa in "NT":
                for diag in "NU":
                    for trial in range(4):
                        m, n = int(rng.integers(2, 6)), int(rng.integers(2, 6))
                        na = m if side == "L" else n
                        # trial 0 dense; 1 zeros in B; 2 and 3: a NaR that only a skip
                        # keeps out of the result (left: in A's used triangle, met by
                        # settled zeros of B; right: in B, met by zero multipliers of A)
                        A = _tri_matrix(rng, na, uplo, nar_off=(trial >= 2 and side == "L"))
                        B = _matrix(rng, m, n, -2.0, 2.0, zeros=0.3 if trial else 0.0,
                                    nars=0.15 if (trial >= 2 and side == "R") else 0.0)
                        if side == "R" and trial == 3:
                            # the first settled column of B is NaR in every row, and every
                            # other multiplier that would carry it onwards is zero
                            k0 = 0 if (uplo == "U") == (transa == "N") else n - 1
                            for i in range(m):
                                B[i][k0] = NAR
                            for j in range(n):
                                if j != k0 and j % 2 == 0:
                                    if transa == "N":
                                        A[k0][j] = ZERO
                                    else:
                                        A[j][k0] = ZERO
                        for alpha in ALPHAS:
                            want = trsm_exact(side, uplo, transa, diag, alpha, A, B)
                            got = _np(B)
                            rc = oracle.rtrsm(side, uplo, transa, diag, m, n, alpha, _np(A), got)
                            if rc:
                                fails.append(f"trsm {side}{uplo}{transa}{diag}: returned {rc}")
                                continue
                            fails += _diff(f"trsm {side}{uplo}{transa}{diag} alpha={alpha:#x} trial {trial}", got,
                                           want)
    return fails


# ---------------------------------------------------------------------------------------
# Rgetrs (Rgetrs.cpp:86-113) and Rpotrs (Rpotrs.cpp:86-105)
# ---------------------------------------------------------------------------------------
def _swap_rows(B, ipiv, forward):
    """Row interchanges of Rlaswp on all columns: i-th interchange swaps rows i and
    ipiv[i] (1-based), applied for i ascending (forward) or descending (inverse)."""
    B = [row[:] for row in B]
    idx = range(len(ipiv)) if forward else range(len(ipiv) - 1, -1, -1)
    for i in idx:
        p = ipiv[i] - 1
        B[i], B[p] = B[p], B[i]
    return B


def getrs_exact(trans, F, ipiv, B):
    if trans == "N":  # P*L*U*x = b: interchanges, L (unit), U
        Y = _swap_rows(B, ipiv, True)
        Y = trsm_exact("L", "L", "N", "U", ONE, F, Y)
        return trsm_exact("L", "U", "N", "N", ONE, F, Y)
    Y = trsm_exact("L", "U", "T", "N", ONE, F, B)  # U^T, then L^T (unit), interchanges backwards
    Y = trsm_exact("L", "L", "T", "U", ONE, F, Y)
    return _swap_rows(Y, ipiv, False)


def potrs_exact(uplo, F, B):
    if uplo == "U":  # U^T U x = b
        Y = trsm_exact("L", "U", "T", "N", ONE, F, B)
        return trsm_exact("L", "U", "N", "N", ONE, F, Y)
    Y = trsm_exact("L", "L", "N", "N", ONE, F, B)  # L L^T x = b
    return trsm_exact("L", "L", "T", "N", ONE, F, Y)


def check_getrs(oracle, seed=2):
    rng = np.random.default_rng(seed)
    fails = []
    for trans in "NT":
        for trial in range(4):
            n, nrhs = int(rng.integers(2, 7)), int(rng.integers(1, 4))
            # factors as Rgetrf leaves them: unit-lower L below, U on and above the
            # diagonal, and a pivot vector with i <= ipiv[i] <= n
            F = _matrix(rng, n, n, -1.0, 1.0, zeros=0.2)
            for i in range(n):
                F[i][i] = X.from_double(float(rng.uniform(1.0, 2.0)))
            ipiv = [int(rng.integers(i + 1, n + 1)) for i in range(n)]
            B = _matrix(rng, n, nrhs, -2.0, 2.0, zeros=0.2 if trial % 2 else 0.0)
            want = getrs_exact(trans, F, ipiv, B)
            got = _np(B)
            rc, info = oracle.rgetrs(trans, _np(F), np.array(ipiv, np.int32), got)
            if rc or info:
                fails.append(f"getrs {trans}: returned {rc}, info {info}")
                continue
            fails += _diff(f"getrs {trans} trial {trial}", got, want)
    return fails


def check_potrs(oracle, seed=3):
    rng = np.random.default_rng(seed)
    fails = []
    for uplo in "UL":
        for trial in range(4):
            n, nrhs = int(rng.integers(2, 7)), int(rng.integers(1, 4))
            F = _matrix(rng, n, n, -1.0, 1.0, zeros=0.2)
            for i in range(n):
                F[i][i] = X.from_double(float(rng.uniform(1.0, 2.0)))
            B = _matrix(rng, n, nrhs, -2.0, 2.0, zeros=0.2 if trial % 2 else 0.0)
            want = potrs_exact(uplo, F, B)
            got = _np(B)
            rc, info = oracle.rpotrs(uplo, _np(F), got)
            if rc or info:
                fails.append(f"potrs {uplo}: returned {rc}, info {info}")
                continue
            fails += _diff(f"potrs {uplo} trial {trial}", got, want)
    return fails


# ---------------------------------------------------------------------------------------
# Rgemv (Rgemv.cpp:116-173)
# ---------------------------------------------------------------------------------------
def _logical(v, length, inc):
    """Element j of a strided vector: v[j*inc] for inc > 0, v[(length-1-j)*|inc|] for
    inc < 0 (the reference starts at 1 - (len-1)*inc, :97-107)."""
    return [v[j * inc] if inc > 0 else v[(length - 1 - j) * (-inc)] for j in range(length)]


def gemv_exact(trans, alpha, A, x, incx, beta, y, incy):
    m, n = len(A), len(A[0])
    lenx, leny = (n, m) if trans == "N" else (m, n)
    y = list(y)
    if m == 0 or n == 0 or (alpha == ZERO and beta == ONE):
        return y
    xs = _logical(x, lenx, incx)
    pos = [j * incy if incy > 0 else (leny - 1 - j) * (-incy) for j in range(leny)]
    ys = [y[p] for p in pos]
    if beta != ONE:  # :116-141
        ys = [ZERO if beta == ZERO else X.mul(beta, v) for v in ys]
    if alpha != ZERO:
        if trans == "N":  # y_i += (alpha * x_j) * a_ij for j ascending (:150-162)
            for i in range(m):
                for j in range(n):
                    ys[i] = X.add(ys[i], X.mul(X.mul(alpha, xs[j]), A[i][j]))
        else:  # y_j += alpha * (sum of a_ij * x_i from zero, i ascending) (:163-172)
            for j in range(n):
                t = ZERO
                for i in range(m):
                    t = X.add(t, X.mul(A[i][j], xs[i]))
                ys[j] = X.add(ys[j], X.mul(alpha, t))
    for p, v in zip(pos, ys):
        y[p] = v
    return y


def check_gemv(oracle, seed=4):
    rng = np.random.default_rng(seed)
    fails = []
    for trans in "NT":
        for incx, incy in ((1, 1), (2, 1), (-1, 1), (1, -2), (-2, -1), (3, 2)):
            for alpha, beta in ((ONE, ZERO), (0x3A000000, MONE), (ZERO, 0x44000000), (MONE, ONE)):
                m, n = int(rng.integers(1, 6)), int(rng.integers(1, 6))
                A = _matrix(rng, m, n, -2.0, 2.0, zeros=0.2)
                lenx, leny = (n, m) if trans == "N" else (m, n)
                x = _pos(rng.uniform(-2, 2, 1 + (lenx - 1) * abs(incx)))
                y = _pos(rng.uniform(-2, 2, 1 + (leny - 1) * abs(incy)))
                if rng.random() < 0.3:  # a NaR in y: beta == 0 must clear it without reading it
                    y[0] = NAR
                want = gemv_exact(trans, alpha, A, x, incx, beta, y, incy)
                got = np.array(y, np.uint32)
                rc = oracle.rgemv(trans, m, n, alpha, _np(A), np.array(x, np.uint32), incx, beta, got, incy)
                if rc:
                    fails.append(f"gemv {trans}: returned {rc}")
                    continue
                fails += _diff(f"gemv {trans} incx={incx} incy={incy} alpha={alpha:#x} beta={beta:#x}", got, want)
    return fails


# ---------------------------------------------------------------------------------------
# Rsyrk (Rsyrk.cpp:96-210)
# ---------------------------------------------------------------------------------------
def syrk_exact(uplo, trans, alpha, A, beta, C):
    n = len(C)
    k = len(A[0]) if trans == "N" else len(A)
    C = [row[:] for row in C]
    tri = [(i, j) for j in range(n) for i in range(n) if (i <= j if uplo == "U" else i >= j)]
    if n == 0 or ((alpha == ZERO or k == 0) and beta == ONE):
        return C
    if alpha == ZERO:  # :96-127
        for i, j in tri:
            C[i][j] = ZERO if beta == ZERO else X.mul(beta, C[i][j])
        return C
    for i, j in tri:
        if trans == "N":
            # beta-op, then c_ij += (alpha * a_jl) * a_il for l ascending, skipping
            # a_jl == 0 (:158-176)
            c = ZERO if beta == ZERO else (C[i][j] if beta == ONE else X.mul(beta, C[i][j]))
            for l in range(k):
                if A[j][l] != ZERO:
                    c = X.add(c, X.mul(X.mul(alpha, A[j][l]), A[i][l]))
        else:
            # sum of a_li * a_lj from zero, then alpha*t (+ beta*c) (:182-210)
            t = ZERO
            for l in range(k):
                t = X.add(t, X.mul(A[l][i], A[l][j]))
            c = X.mul(alpha, t) if beta == ZERO else X.add(X.mul(alpha, t), X.mul(beta, C[i][j]))
        C[i][j] = c
    return C


def check_syrk(oracle, seed=5):
    rng = np.random.default_rng(seed)
    fails = []
    for uplo in "UL":
        for trans in "NT":
            for alpha, beta in ((ONE, ZERO), (MONE, ONE), (0x3A000000, 0xC8000000), (ZERO, 0x44000000)):
                for trial in range(3):
                    n, k = int(rng.integers(2, 6)), int(rng.integers(1, 6))
                    A = _matrix(rng, n if trans == "N" else k, k if trans == "N" else n, -2.0, 2.0,
                                zeros=0.3, nars=0.05 if trial == 1 else 0.0)
                    if trial == 2 and trans == "N":
                        # a zero multiplier a_jl meets a NaR a_il of the same column l:
                        # only the skip (:149, :169) keeps element (i, j) finite
                        j0, i0 = (n - 1, 0) if uplo == "U" else (0, n - 1)
                        A[j0][0], A[i0][0] = ZERO, NAR
                    C = _matrix(rng, n, n, -2.0, 2.0)
                    want = syrk_exact(uplo, trans, alpha, A, beta, C)
                    got = _np(C)
                    rc = oracle.rsyrk(uplo, trans, n, k, alpha, _np(A), beta, got)
                    if rc:
                        fails.append(f"syrk {uplo}{trans}: returned {rc}")
                        continue
                    fails += _diff(f"syrk {uplo}{trans} alpha={alpha:#x} beta={beta:#x} trial {trial}", got, want)
    return fails


# ---------------------------------------------------------------------------------------
# Upper Cholesky (Rpotrf_opencl.cpp:100-124 / Rpotrf.cpp, Rpotrf2.cpp:86-134)
# ---------------------------------------------------------------------------------------
def _le_zero(p):
    return p == 0 or (p >> 31) == 1  # signed pattern <= 0 (NaR included)


def potrf_upper_exact(A, nb):
    """Upper Cholesky element by element.  Top level, block J = [j0, j0 + jb): every
    element (i, c) with i in J and c >= i first gets  -sum_{l < j0} u_li * u_lc  (one sum
    from zero, ascending l, then added: the Rsyrk 'U','T' / RgemmG 'T','N' update with
    alpha = -1, beta = 1).  The diagonal block is then factored by the halving recursion of
    Rpotrf2: left half, then its rows of the right half solved (dot form, ascending k,
    then divided by the diagonal), then the right half's triangle gets -sum over the
    left half's rows (one sum from zero), then the right half.  Rows of J right of the
    block are solved the same way against the factored block (Rtrsm 'L','U','T','N').
    Returns (U, info)."""
    n = len(A)
    U = [row[:] for row in A]
    if nb <= 0:
        nb = 64

    def group(i, c, lo, hi):
        """u_ic += -(sum of u_li u_lc, l in [lo, hi), from zero) when the range is not empty."""
        if hi <= lo:
            return
        t = ZERO
        for l in range(lo, hi):
            t = X.add(t, X.mul(U[l][i], U[l][c]))
        U[i][c] = X.add(neg(t), U[i][c])

    def solve_rows(lo, hi, cols):
        """Rows [lo, hi) of columns `cols` against the factored block [lo, hi)."""
        for c in cols:
            for i in range(lo, hi):
                t = U[i][c]
                for k in range(lo, i):
                    t = sub(t, X.mul(U[k][i], U[k][c]))
                U[i][c] = X.div(t, U[i][i])

    def rec(o, s):
        if s == 1:
            if _le_zero(U[o][o]):
                return 1
            U[o][o] = X.sqrt(U[o][o])
            return 0
        n1 = s // 2
        f = rec(o, n1)
        if f:
            return f
        solve_rows(o, o + n1, range(o + n1, o + s))
        for c in range(o + n1, o + s):
            for i in range(o + n1, c + 1):
                group(i, c, o, o + n1)
        f = rec(o + n1, s - n1)
        return f + n1 if f else 0

    if nb <= 1 or nb >= n:
        return U, rec(0, n)
    for j0 in range(0, n, nb):
        jb = min(nb, n - j0)
        for c in range(j0, j0 + jb):  # the diagonal block (Rsyrk 'U','T')
            for i in range(j0, c + 1):
                group(i, c, 0, j0)
        f = rec(j0, jb)
        if f:
            return U, f + j0
        for c in range(j0 + jb, n):  # the block's rows right of it (RgemmG 'T','N')
            for i in range(j0, j0 + jb):
                group(i, c, 0, j0)
        solve_rows(j0, j0 + jb, range(j0 + jb, n))
    return U, 0


def _spd_exact(rng, n):
    M = rng.uniform(-1, 1, (n, n))
    S = M @ M.T / n + np.eye(n)
    return [[X.from_double(float(S[i, j])) for j in range(n)] for i in range(n)]


def check_potrf_upper(oracle, seed=6):
    rng = np.random.default_rng(seed)
    fails = []
    for n, nb in ((6, 0), (7, 1), (7, 2), (7, 3), (5, 2)):
        A = _spd_exact(rng, n)
        U, info = potrf_upper_exact(A, nb)
        got = _np(A)
        rc, oinfo = oracle.rpotrf("U", got, nb=nb, nthreads=1)
        if rc or oinfo != info:
            fails.append(f"potrf U n={n} nb={nb}: info {oinfo}, exact {info}")
            continue
        want = np.array(A, np.uint32)
        for i in range(n):
            for c in range(i, n):
                want[i, c] = U[i][c]
        fails += _diff(f"potrf U n={n} nb={nb}", got, want)
    # a failing leading minor: the same info from both
    A = _spd_exact(rng, 6)
    A[3][3] = X.from_double(-1.0)
    for nb in (0, 2):
        _, info = potrf_upper_exact(A, nb)
        rc, oinfo = oracle.rpotrf("U", _np(A), nb=nb, nthreads=1)
        if oinfo != info or info != 4:
            fails.append(f"potrf U not PD nb={nb}: info {oinfo}, exact {info}")
    return fails


# ---------------------------------------------------------------------------------------
# iRamax (iRamax.cpp:31-88, abs util.cpp:12-17)
# ---------------------------------------------------------------------------------------
def check_iramax(oracle, seed=7):
    rng = np.random.default_rng(seed)
    fails = []
    for trial in range(40):
        n = int(rng.integers(1, 9))
        inc = int(rng.integers(1, 3))
        vals = rng.choice([-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0], n)  # many ties in |x|
        x = np.zeros(1 + (n - 1) * inc, np.uint32)
        x[::inc] = _pos(vals)
        mags = [abs(X.decode(int(v))) for v in x[::inc]]
        want = 1 + mags.index(max(mags))  # the first index of the largest magnitude
        got = oracle.iramax(x, inc, n)
        if got != want:
            fails.append(f"iramax {vals.tolist()} inc {inc}: oracle {got}, first maximum {want}")
    return fails


CHECKS = {
    "trsm": check_trsm,
    "getrs": check_getrs,
    "potrs": check_potrs,
    "gemv": check_gemv,
    "syrk": check_syrk,
    "potrf_upper": check_potrf_upper,
    "iramax": check_iramax,
}
