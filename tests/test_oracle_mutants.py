"""The oracle's pins have teeth: a copy of the oracle with one plausible mistake planted
(a reciprocal for a quotient, a sweep in the wrong direction, a zero skip dropped, a sum
in the wrong order, an interchange direction flipped, alpha folded into a term, ...) is
built in a temporary directory and must fail at least one independent check
(tests/oracle_pins.py, or the exact-rational brute-force tests of test_oracle_blas.py).
The product is not involved; CPU only."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
RP = "rpack_oracle.c"

# name, file, text to replace, occurrence (1-based), replacement, checks that must fail
# ("blas" = the brute-force subset of test_oracle_blas.py)
MUTANTS = [
    ("trsm LUNN reciprocal for quotient", RP, "B_(kk, j) = DIV(B_(kk, j), A_(kk, kk));", 1,
     "B_(kk, j) = MUL(DIV(ONE, A_(kk, kk)), B_(kk, j));", ["trsm", "getrs", "potrs"]),
    ("trsm LLNN reciprocal for quotient", RP, "B_(kk, j) = DIV(B_(kk, j), A_(kk, kk));", 2,
     "B_(kk, j) = MUL(DIV(ONE, A_(kk, kk)), B_(kk, j));", ["trsm", "potrs"]),
    ("trsm LUNN zero skip dropped", RP, "if (B_(kk, j) != ZERO) {", 1, "if (1) {", ["trsm"]),
    ("trsm LLNN zero skip dropped", RP, "if (B_(kk, j) != ZERO) {", 2, "if (1) {", ["trsm"]),
    ("trsm LUNN swept top-down", RP, "for (int kk = m - 1; kk >= 0; kk--) {", 1, "for (int kk = 0; kk < m; kk++) {",
     ["trsm", "getrs"]),
    ("trsm LUTN terms in descending k", RP, "for (int kk = 0; kk < i; kk++)", 1, "for (int kk = i - 1; kk >= 0; kk--)",
     ["trsm", "getrs", "potrs"]),
    ("trsm LLTN terms in descending k", RP, "for (int kk = i + 1; kk < m; kk++)", 1,
     "for (int kk = m - 1; kk > i; kk--)", ["trsm", "potrs"]),
    ("trsm LLTN reciprocal for quotient", RP, "temp = DIV(temp, A_(i, i));", 2,
     "temp = MUL(DIV(ONE, A_(i, i)), temp);", ["trsm", "potrs"]),
    ("trsm RUNN zero skip dropped", RP, "if (A_(kk, j) != ZERO)", 1, "if (1)", ["trsm"]),
    ("trsm RLNN quotient for reciprocal", RP, "B_(i, j) = MUL(temp, B_(i, j));", 2,
     "B_(i, j) = DIV(B_(i, j), A_(j, j));", ["trsm"]),
    ("trsm RUTN alpha forgotten", RP, "B_(i, kk) = MUL(alpha, B_(i, kk));", 1, "B_(i, kk) = MUL(ONE, B_(i, kk));",
     ["trsm"]),
    ("trsm RLTN quotient for reciprocal", RP, "B_(i, kk) = MUL(temp, B_(i, kk));", 2,
     "B_(i, kk) = DIV(B_(i, kk), A_(kk, kk));", ["trsm"]),
    ("trsm RLTN zero skip dropped", RP, "if (A_(j, kk) != ZERO) {", 2, "if (1) {", ["trsm"]),
    ("gemv N alpha folded into each term", RP, "y[iy] = ADD(y[iy], MUL(temp, a[i + (size_t)j * lda]));", 1,
     "y[iy] = ADD(y[iy], MUL(alpha, MUL(x[jx], a[i + (size_t)j * lda])));", ["gemv"]),
    ("gemv negative incx read forwards", RP, "int kx = (incx > 0) ? 0 : -(lenx - 1) * incx;", 1,
     "int kx = 0;\n    if (incx < 0) incx = -incx;", ["gemv"]),
    ("gemv T alpha in every product", RP, "temp = ADD(temp, MUL(a[i + (size_t)j * lda], x[ix]));", 1,
     "temp = ADD(temp, MUL(MUL(alpha, a[i + (size_t)j * lda]), x[ix]));", ["gemv"]),
    ("gemv beta 0 multiplies", RP, "y[iy] = (beta == ZERO) ? ZERO : MUL(beta, y[iy]);", 1,
     "y[iy] = MUL(beta, y[iy]);", ["gemv"]),
    ("syrk N zero skip dropped", RP, "if (A_(j, l) != ZERO) {", 1, "if (1) {", ["syrk"]),
    ("syrk N alpha folded into each term", RP, "C_(i, j) = ADD(C_(i, j), MUL(temp, A_(i, l)));", 1,
     "C_(i, j) = ADD(C_(i, j), MUL(alpha, MUL(A_(j, l), A_(i, l))));", ["syrk"]),
    ("syrk T sum in descending l", RP,
     "for (int l = 0; l < k; l++)\n                    temp = ADD(temp, MUL(A_(l, i), A_(l, j)));", 1,
     "for (int l = k - 1; l >= 0; l--)\n                    temp = ADD(temp, MUL(A_(l, i), A_(l, j)));",
     ["syrk", "potrf_upper"]),
    ("getrs T interchanges applied forwards", RP, "orc_rlaswp(nrhs, b, ldb, 1, n, ipiv, -1);", 1,
     "orc_rlaswp(nrhs, b, ldb, 1, n, ipiv, 1);", ["getrs"]),
    ("getrs N solves in the wrong order", RP,
     "orc_rtrsm('L', 'L', 'N', 'U', n, nrhs, ONE, a, lda, b, ldb);\n"
     "        orc_rtrsm('L', 'U', 'N', 'N', n, nrhs, ONE, a, lda, b, ldb);", 1,
     "orc_rtrsm('L', 'U', 'N', 'N', n, nrhs, ONE, a, lda, b, ldb);\n"
     "        orc_rtrsm('L', 'L', 'N', 'U', n, nrhs, ONE, a, lda, b, ldb);", ["getrs"]),
    ("potrs L second solve with a unit diagonal", RP, "orc_rtrsm('L', 'L', 'T', 'N', n, nrhs, ONE, a, lda, b, ldb);",
     1, "orc_rtrsm('L', 'L', 'T', 'U', n, nrhs, ONE, a, lda, b, ldb);", ["potrs"]),
    ("potrf2 splits at the upper half", RP, "int n1 = n / 2, n2 = n - n1, iinfo = 0;", 1,
     "int n1 = (n + 1) / 2, n2 = n - n1, iinfo = 0;", ["potrf_upper"]),
    ("potrf U panel solve with a unit diagonal", RP,
     "orc_rtrsm('L', 'U', 'T', 'N', jb, n - j - jb + 1, ONE, A_(j - 1, j - 1), lda,", 1,
     "orc_rtrsm('L', 'U', 'T', 'U', jb, n - j - jb + 1, ONE, A_(j - 1, j - 1), lda,", ["potrf_upper"]),
    ("iramax keeps the last maximum", RP, "if (p_gt(v, dmax)) {", 1, "if (!p_gt(dmax, v)) {", ["iramax"]),
    ("gemm NN alpha folded into each term", RP, "cj[i] = ADD(cj[i], MUL(temp, al[i]));", 1,
     "cj[i] = ADD(cj[i], MUL(alpha, MUL(blj, al[i])));", ["blas"]),
    ("gemm NN beta ignored", RP, "cj[i] = MUL(beta, cj[i]);", 1, "cj[i] = cj[i];", ["blas"]),
    ("getrf2 quotient for the reciprocal", RP, "orc_rscal(m - 1, DIV(ONE, a[0]), a + 1, 1);", 1,
     "for (int r = 1; r < m; r++) a[r] = DIV(a[r], a[0]);", ["blas"]),
    ("potrf2 L panel solve with a unit diagonal", RP,
     "orc_rtrsm('R', 'L', 'T', 'N', n2, n1, ONE, a, lda, a + n1, lda);", 1,
     "orc_rtrsm('R', 'L', 'T', 'U', n2, n1, ONE, a, lda, a + n1, lda);", ["blas"]),
]

BLAS_SUBSET = "brute_force or ascending or identity_and_special or singular_info or not_pd"

_PROBE = """
import sys
sys.path[:0] = [{tests!r}, {golden!r}]
from oracle_lib import Oracle
from oracle_pins import CHECKS
o = Oracle(sys.argv[1])
print(sum(len(CHECKS[c](o)) for c in sys.argv[2:]))
"""


def _build(tmp, idx, mutant):
    name, fname, old, occ, new, _ = mutant
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d)
    for f in ("posit32_oracle.c", "posit32_oracle.h", "rpack_oracle.c"):
        shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    src = open(path).read()
    pos = -1
    for _ in range(occ):
        pos = src.find(old, pos + 1)
        assert pos >= 0, f"{name}: text not found (occurrence {occ})"
    src = src[:pos] + new + src[pos + len(old):]
    open(path, "w").write(src)
    lib = os.path.join(d, "libmut.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-std=gnu11", "-shared", "-o", lib,
                    os.path.join(d, "posit32_oracle.c"), os.path.join(d, "rpack_oracle.c"), "-lm"],
                   check=True, capture_output=True)
    return lib


def _caught(lib, checks):
    """Number of failing checks (pins) or whether the brute-force subset failed (blas)."""
    out = []
    pins = [c for c in checks if c != "blas"]
    if pins:
        r = subprocess.run([sys.executable, "-c", _PROBE.format(tests=HERE, golden=os.path.join(HERE, "golden")),
                            lib, *pins], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        out.append(int(r.stdout.split()[-1]) > 0)
    if "blas" in checks:
        env = dict(os.environ, PGA_ORACLE_LIB=lib)
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                            os.path.join(HERE, "test_oracle_blas.py"), "-k", BLAS_SUBSET],
                           env=env, capture_output=True, text=True, timeout=600)
        out.append(r.returncode == 1)  # 1: tests ran and some failed
    return all(out)


@pytest.fixture(scope="module")
def verdicts(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))

    def one(i):
        lib = _build(tmp, i, MUTANTS[i])
        return _caught(lib, MUTANTS[i][5])

    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, range(len(MUTANTS))))


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(verdicts, idx):
    assert verdicts[idx], f"the checks {MUTANTS[idx][5]} did not catch: {MUTANTS[idx][0]}"
