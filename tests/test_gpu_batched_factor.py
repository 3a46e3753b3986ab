"""Batched factorizations and solves (one persistent launch): every task must equal the
per-call device routine bit for bit (factors, pivots, info, solutions) and, where the
oracle is cheap, the oracle's restatement of the reference routine."""
import os

import numpy as np
import pytest

from datagen import posit_matrix, uniform_doubles

pytestmark = pytest.mark.gpu

ONE, NAR = 0x40000000, 0x80000000
THREADS = min(16, os.cpu_count() or 1)


def _spd(oracle, seed, n):
    M = uniform_doubles(seed, n * n).reshape(n, n)
    S = M @ M.T / n + np.eye(n)
    return np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))


def _same(got, want, what):
    bad = np.argwhere(np.asarray(got) != np.asarray(want))
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {bad[:4].tolist()}"


class Case:
    """One task: its inputs, the device buffers the plan works on, and a per-call twin."""

    def __init__(self, pga, kind, A, uplo="L", nb=0, B=None, factored=None):
        self.pga, self.kind, self.uplo, self.nb = pga, kind, uplo, nb
        self.n = A.shape[0]
        self.A = A.copy(order="F")
        self.B = None if B is None else B.copy(order="F")
        self.factored = factored  # (factors, ipiv) for solve-only tasks

    def buffers(self):
        pga, n = self.pga, self.n
        src = self.A if self.factored is None else self.factored[0]
        d = {"a": pga.DeviceArray.from_host(src.copy(order="F")), "info": pga.DeviceArray(1, np.int32, "C")}
        d["info"].copy_from_host(np.array([-7], np.int32))
        if self.kind in ("getrf", "gesv", "getrs"):
            piv = np.zeros(max(1, n), np.int32) if self.factored is None else self.factored[1]
            d["ipiv"] = pga.DeviceArray.from_host(piv.copy())
        if self.B is not None:
            d["b"] = pga.DeviceArray.from_host(self.B.copy(order="F"))
        return d

    def task(self, d):
        t = {"kind": self.kind, "n": self.n, "lda": max(1, self.n), "a": d["a"], "info": d["info"], "uplo": self.uplo,
             "nb": self.nb}
        if "ipiv" in d:
            t["ipiv"] = d["ipiv"]
        if "b" in d:
            t.update(b=d["b"], ldb=max(1, self.n), nrhs=self.B.shape[1])
        return t

    def per_call(self):
        """The same work through the per-call device routines."""
        pga, n = self.pga, self.n
        d = self.buffers()
        info = 0
        if self.kind in ("getrf", "gesv"):
            pga.getrf_device(n, n, d["a"], max(1, n), d["ipiv"], d["info"], self.nb)
        if self.kind in ("potrf", "posv"):
            pga.potrf_device(self.uplo, n, d["a"], max(1, n), d["info"], self.nb)
        if self.kind in ("getrf", "gesv", "potrf", "posv"):
            pga.synchronize()
            info = int(d["info"].to_host()[0])
        else:
            d["info"].copy_from_host(np.array([0], np.int32))
        if self.B is not None and info == 0:
            if self.kind in ("gesv", "getrs"):
                pga.getrs_device("N", n, self.B.shape[1], d["a"], max(1, n), d["ipiv"], d["b"], max(1, n))
            else:
                pga.potrs_device(self.uplo, n, self.B.shape[1], d["a"], max(1, n), d["b"], max(1, n))
        pga.synchronize()
        return self.read(d)

    def read(self, d):
        out = {"a": d["a"].to_host(), "info": int(d["info"].to_host()[0])}
        if "ipiv" in d:
            out["ipiv"] = d["ipiv"].to_host()
        if "b" in d:
            out["b"] = d["b"].to_host()
        return out


def _check_plan(pga, cases, graph=False):
    bufs = [c.buffers() for c in cases]
    plan = pga.FactorPlan([c.task(d) for c, d in zip(cases, bufs)])
    if graph:
        plan.capture_graph()
        plan.launch_graph()
    else:
        plan.execute()
    pga.synchronize()
    for i, (c, d) in enumerate(zip(cases, bufs)):
        got, want = c.read(d), c.per_call()
        what = f"task {i} {c.kind} n={c.n} uplo={c.uplo} nb={c.nb}"
        assert got["info"] == want["info"], (what, got["info"], want["info"])
        for key in want:
            if key != "info":
                _same(got[key], want[key], f"{what} {key}")
    return plan, bufs


def test_mixed_tasks_equal_per_call(pga, oracle):
    rng = np.random.default_rng(11)
    cases = []
    for i in range(24):
        n = int(rng.integers(1, 160))
        kind = ["getrf", "potrf", "gesv", "posv"][i % 4]
        uplo = "LU"[(i // 4) % 2]
        nb = [0, 16, 64, 1][i % 4]
        A = posit_matrix(oracle, 300 + i, n, n) if kind in ("getrf", "gesv") else _spd(oracle, 300 + i, n)
        B = posit_matrix(oracle, 400 + i, n, int(rng.integers(1, 5))) if kind in ("gesv", "posv") else None
        cases.append(Case(pga, kind, A, uplo, nb, B))
    _check_plan(pga, cases)


def test_solve_only_tasks(pga, oracle):
    cases = []
    for i, n in enumerate([5, 64, 130]):
        A = posit_matrix(oracle, 500 + i, n, n)
        F = A.copy(order="F")
        _, piv, info = oracle.rgetrf(F, nb=0, nthreads=THREADS)
        cases.append(Case(pga, "getrs", A, B=posit_matrix(oracle, 510 + i, n, 3), factored=(F, piv)))
        for uplo in "LU":
            S = _spd(oracle, 520 + i, n)
            G = S.copy(order="F")
            assert oracle.rpotrf(uplo, G, nb=0, nthreads=THREADS)[1] == 0
            cases.append(Case(pga, "potrs", S, uplo, B=posit_matrix(oracle, 530 + i, n, 2), factored=(G, None)))
    _check_plan(pga, cases)


def test_special_matrices(pga, oracle):
    """Zero pivots, NaR entries, rank deficiency, a failing Cholesky step: the reference
    leaves structure-dependent bits here (skipped zero multipliers, partial factors)."""
    n = 96
    cases = []
    for nb in (0, 8, 32):
        Z = posit_matrix(oracle, 600, n, n)
        Z[:, 40] = 0
        N = posit_matrix(oracle, 601, n, n)
        N[7, 3] = NAR
        N[50, 60] = NAR
        N[20, 9] = 0
        R = posit_matrix(oracle, 602, n, n)
        R[:, 10] = R[:, 20]
        for M in (Z, N, R):
            cases.append(Case(pga, "getrf", M, nb=nb))
        for uplo in "LU":
            P = _spd(oracle, 603, n)
            P[61, 61] = oracle.from_double(-2.0)
            cases.append(Case(pga, "posv", P, uplo, nb, B=posit_matrix(oracle, 604, n, 2)))
            Q = _spd(oracle, 605, n)
            Q[30, 5] = NAR
            Q[5, 30] = NAR
            Q[44, 2] = 0
            Q[2, 44] = 0
            cases.append(Case(pga, "potrf", Q, uplo, nb))
    _check_plan(pga, cases)


def test_against_oracle(pga, oracle):
    """Small problems checked directly against the oracle as well."""
    cases, wants = [], []
    for i, n in enumerate([1, 2, 17, 70]):
        A = posit_matrix(oracle, 700 + i, n, n)
        F = A.copy(order="F")
        _, piv, info = oracle.rgetrf(F, nb=16, nthreads=THREADS)
        cases.append(Case(pga, "getrf", A, nb=16))
        wants.append(("lu", F, piv, info))
        S = _spd(oracle, 710 + i, n)
        G = S.copy(order="F")
        _, info = oracle.rpotrf("L", G, nb=16, nthreads=THREADS)
        cases.append(Case(pga, "potrf", S, "L", 16))
        wants.append(("chol", G, None, info))
    bufs = [c.buffers() for c in cases]
    plan = pga.FactorPlan([c.task(d) for c, d in zip(cases, bufs)])
    plan.execute()
    pga.synchronize()
    for c, d, (kind, F, piv, info) in zip(cases, bufs, wants):
        got = c.read(d)
        assert got["info"] == info
        if kind == "lu":
            _same(got["ipiv"], piv, "ipiv")
            _same(got["a"], F, "LU factors")
        else:
            _same(np.tril(got["a"]), np.tril(F), "Cholesky factor")


def test_graph_replays(pga, oracle):
    """Capture once; replay with the inputs restored by eager copies in between."""
    cases = [Case(pga, "gesv", posit_matrix(oracle, 800 + i, n, n), nb=16, B=posit_matrix(oracle, 810 + i, n, 2))
             for i, n in enumerate([33, 100, 257])]
    cases += [Case(pga, "posv", _spd(oracle, 820, 150), "U", 0, B=posit_matrix(oracle, 821, 150, 1))]
    plan, bufs = _check_plan(pga, cases, graph=True)
    for rep in range(2):
        for c, d in zip(cases, bufs):
            d["a"].copy_from_host(c.A.copy(order="F"))
            d["b"].copy_from_host(c.B.copy(order="F"))
        pga.synchronize()
        plan.launch_graph()
        pga.synchronize()
        for i, (c, d) in enumerate(zip(cases, bufs)):
            got, want = c.read(d), c.per_call()
            assert got["info"] == want["info"]
            for key in ("a", "b"):
                _same(got[key], want[key], f"replay {rep} task {i} {key}")


def test_largest_order(pga, oracle):
    n = 512
    cases = [Case(pga, "gesv", posit_matrix(oracle, 900, n, n), B=posit_matrix(oracle, 901, n, 1)),
             Case(pga, "potrf", _spd(oracle, 902, n), "L"), Case(pga, "potrf", _spd(oracle, 903, n), "U")]
    _check_plan(pga, cases)


def test_blocked_tasks_equal_per_call(pga, oracle):
    """Tasks larger than one workgroup handles advance together one column block at a
    time; every kind, both triangles, several block sizes (1 and >= n: unblocked), mixed
    with one-workgroup tasks in the same plan."""
    rng = np.random.default_rng(21)
    cases = []
    specs = [("gesv", 700, "L", 0), ("posv", 600, "U", 0), ("posv", 530, "L", 64), ("getrf", 513, "L", 100),
             ("potrf", 777, "U", 48), ("potrf", 520, "L", 0), ("gesv", 530, "L", 1), ("posv", 515, "U", 515),
             ("gesv", 90, "L", 16), ("posv", 200, "U", 0)]
    for i, (kind, n, uplo, nb) in enumerate(specs):
        A = posit_matrix(oracle, 1000 + i, n, n) if kind in ("getrf", "gesv") else _spd(oracle, 1000 + i, n)
        B = posit_matrix(oracle, 1100 + i, n, int(rng.integers(1, 4))) if kind in ("gesv", "posv") else None
        cases.append(Case(pga, kind, A, uplo, nb, B))
    _check_plan(pga, cases)


def test_blocked_solve_only_and_failures(pga, oracle):
    """Solve-only tasks on given factors; a zero column (LU info, factorization goes on)
    and a failing Cholesky block (info, every later step skipped) at sizes of the
    blocked path."""
    n = 600
    cases = []
    A = posit_matrix(oracle, 1200, n, n)
    F = pga.DeviceArray.from_host(A.copy(order="F"))
    ip = pga.DeviceArray(n, np.int32, "C")
    inf = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(n, n, F, n, ip, inf, 64)
    pga.synchronize()
    cases.append(Case(pga, "getrs", A, B=posit_matrix(oracle, 1201, n, 2), factored=(F.to_host(), ip.to_host())))
    for uplo in "LU":
        S = _spd(oracle, 1202, n)
        G = pga.DeviceArray.from_host(S.copy(order="F"))
        pga.potrf_device(uplo, n, G, n, inf, 64)
        pga.synchronize()
        cases.append(Case(pga, "potrs", S, uplo, B=posit_matrix(oracle, 1203, n, 2), factored=(G.to_host(), None)))
    Z = posit_matrix(oracle, 1204, n, n)
    Z[:, 300] = 0
    Z[17, 5] = NAR
    cases.append(Case(pga, "gesv", Z, nb=64, B=posit_matrix(oracle, 1205, n, 1)))
    for uplo, bad in (("U", 200), ("L", 450), ("U", 64)):
        P = _spd(oracle, 1206, n)
        P[bad, bad] = oracle.from_double(-2.0)
        cases.append(Case(pga, "posv", P, uplo, 64, B=posit_matrix(oracle, 1207, n, 2)))
    _check_plan(pga, cases)


def test_blocked_solves_registers_and_fallback(pga, oracle):
    """The blocked path's solves hold a right-hand side in FP64 registers, 1024 rows per
    register slot, and redo a sweep at the pattern level when a value leaves the FP64
    range: orders over several slots, ends that are not whole 64-row blocks, and columns
    with a NaR, a huge value (2^100: products out of range) and all zeros."""
    cases = []
    for i, (kind, n, uplo) in enumerate((("getrs", 2100, "L"), ("potrs", 1500, "U"), ("potrs", 1100, "L"),
                                         ("getrs", 1025, "L"), ("potrs", 650, "U"))):
        A = posit_matrix(oracle, 1400 + i, n, n) if kind == "getrs" else _spd(oracle, 1400 + i, n)
        F = pga.DeviceArray.from_host(A.copy(order="F"))
        inf = pga.DeviceArray(1, np.int32, "C")
        ip = pga.DeviceArray(n, np.int32, "C")
        if kind == "getrs":
            pga.getrf_device(n, n, F, n, ip, inf, 64)
        else:
            pga.potrf_device(uplo, n, F, n, inf, 64)
        pga.synchronize()
        assert int(inf.to_host()[0]) == 0
        B = posit_matrix(oracle, 1450 + i, n, 4)
        B[n // 3, 1] = NAR
        B[n - 2, 2] = oracle.from_double(2.0 ** 100)
        B[:, 3] = 0
        cases.append(Case(pga, kind, A, uplo, B=B, factored=(F.to_host(), ip.to_host() if kind == "getrs" else None)))
    _check_plan(pga, cases)


def test_blocked_graph_replays(pga, oracle):
    cases = [Case(pga, "gesv", posit_matrix(oracle, 1300, 640, 640), nb=0, B=posit_matrix(oracle, 1301, 640, 2)),
             Case(pga, "posv", _spd(oracle, 1302, 560), "U", 0, B=posit_matrix(oracle, 1303, 560, 1))]
    plan, bufs = _check_plan(pga, cases, graph=True)
    for rep in range(2):
        for c, d in zip(cases, bufs):
            d["a"].copy_from_host(c.A.copy(order="F"))
            d["b"].copy_from_host(c.B.copy(order="F"))
        pga.synchronize()
        plan.launch_graph()
        pga.synchronize()
        for i, (c, d) in enumerate(zip(cases, bufs)):
            got, want = c.read(d), c.per_call()
            assert got["info"] == want["info"]
            for key in ("a", "b"):
                _same(got[key], want[key], f"replay {rep} task {i} {key}")


def test_autotuned_plan_same_bits(pga, oracle):
    """The plan's tuner (workgroups per CU, task order) times scratch copies and leaves
    the caller's operands alone; the tuned plan, eager and captured, and a plan tuned twice
    give the per-call bits."""
    rng = np.random.default_rng(31)
    cases = []
    for i in range(40):
        n = int(rng.integers(1, 200))
        kind = ["getrf", "potrf", "gesv", "posv", "getrs", "potrs"][i % 6]
        uplo = "LU"[(i // 6) % 2]
        A = posit_matrix(oracle, 1400 + i, n, n) if kind in ("getrf", "gesv", "getrs") else _spd(oracle, 1400 + i, n)
        B = posit_matrix(oracle, 1500 + i, n, int(rng.integers(1, 4))) if kind not in ("getrf", "potrf") else None
        factored = None
        if kind == "getrs":
            F = A.copy(order="F")
            _, piv, _ = oracle.rgetrf(F, nb=0, nthreads=THREADS)
            factored = (F, piv)
        elif kind == "potrs":
            G = A.copy(order="F")
            assert oracle.rpotrf(uplo, G, nb=0, nthreads=THREADS)[1] == 0
            factored = (G, None)
        cases.append(Case(pga, kind, A, uplo, [0, 16, 1][i % 3], B, factored))
    cases.append(Case(pga, "gesv", posit_matrix(oracle, 1600, 600, 600), B=posit_matrix(oracle, 1601, 600, 2)))
    bufs = [c.buffers() for c in cases]
    plan = pga.FactorPlan([c.task(d) for c, d in zip(cases, bufs)])
    plan.capture_graph()
    untouched = [c.read(d) for c, d in zip(cases, bufs)]
    first = plan.autotune()
    second = plan.autotune()
    for t in (first, second):
        assert t["grid"] >= 1 and t["schedule"] in (0, 1, 2) and t["tune_ms"] >= 0, t
    for i, (c, d) in enumerate(zip(cases, bufs)):  # tuning ran on copies
        for key, v in untouched[i].items():
            assert np.array_equal(np.asarray(c.read(d)[key]), np.asarray(v)), (i, key)
    assert plan.info()["grid"] == second["grid"]
    for mode in ("eager", "graph"):
        for c, d in zip(cases, bufs):
            src = c.A if c.factored is None else c.factored[0]
            d["a"].copy_from_host(src.copy(order="F"))
            if "b" in d:
                d["b"].copy_from_host(c.B.copy(order="F"))
            d["info"].copy_from_host(np.array([-7], np.int32))
        pga.synchronize()
        if mode == "eager":
            plan.execute()
        else:
            plan.capture_graph()
            plan.launch_graph()
        pga.synchronize()
        for i, (c, d) in enumerate(zip(cases, bufs)):
            got, want = c.read(d), c.per_call()
            what = f"{mode} task {i} {c.kind} n={c.n} uplo={c.uplo} nb={c.nb}"
            assert got["info"] == want["info"], (what, got["info"], want["info"])
            for key in want:
                if key != "info":
                    _same(got[key], want[key], f"{what} {key}")


def test_rejects_bad_tasks(pga):
    a = pga.DeviceArray((4, 4))
    with pytest.raises(pga.PositError):
        pga.FactorPlan([{"kind": "getrf", "n": 4, "a": a, "lda": 4}])  # no ipiv
    with pytest.raises(pga.PositError):
        pga.FactorPlan([{"kind": "potrf", "n": 20000, "a": a, "lda": 20000, "uplo": "L"}])  # too large
