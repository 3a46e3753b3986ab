"""Serving with continuous batching: every request's result must be the same served in
the mix (batched launches and worker streams overlapping) as alone through the plain
per-call path, bit for bit; admission keeps device memory within the budget; resident
factors serve re-solves and are evicted least-recently-used."""
import os

import numpy as np
import pytest

from datagen import posit_matrix, uniform_doubles

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
MiB = 1 << 20


def _spd(oracle, seed, n):
    M = uniform_doubles(seed, n * n).reshape(n, n)
    S = M @ M.T / n + np.eye(n)
    return np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))


def _requests(oracle, seed, sizes):
    rng = np.random.default_rng(seed)
    reqs = []
    for i, n in enumerate(sizes):
        kind = "lu" if i % 2 == 0 else "chol"
        A = posit_matrix(oracle, seed + 10 * i, n, n) if kind == "lu" else _spd(oracle, seed + 10 * i, n)
        B = posit_matrix(oracle, seed + 10 * i + 1, n, int(rng.integers(1, 5)))
        reqs.append((kind, n, A, B))
    return reqs


def _run(server, reqs, resolves=()):
    """Submit everything at once (plus re-solves of earlier factors); return results."""
    ids = [server.submit(kind, n, A, B) for kind, n, A, B in reqs]
    rids = [server.submit("resolve", reqs[j][1], b=B, factor=ids[j]) for j, B in resolves]
    return [server.wait(i) for i in ids], [server.wait(i) for i in rids]


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_mix_equals_alone(pga, oracle, uplo):
    sizes = [8, 40, 100, 300, 513, 600, 64, 257, 700, 33, 128, 512]
    reqs = _requests(oracle, 100, sizes)
    resolves = [(1, posit_matrix(oracle, 190, 40, 2)), (4, posit_matrix(oracle, 191, 513, 1)),
                (7, posit_matrix(oracle, 192, 257, 3))]
    served = pga.Server(512 * MiB, "batched", batch_max_n=512, batch_slots=2, workers=2, chol_uplo=uplo)
    got, got_r = _run(served, reqs, resolves)
    st = served.stats()
    served.close()
    plain = pga.Server(512 * MiB, "plain", chol_uplo=uplo)
    want, want_r = _run(plain, reqs, resolves)
    plain.close()
    assert st["batches"] >= 1 and st["batched_requests"] >= 1
    for g, w in zip(got + got_r, want + want_r):
        assert g["status"] == w["status"] == pga.REQ_OK
        assert g["info"] == w["info"] == 0
        assert np.array_equal(g["x"], w["x"])


def test_served_matches_oracle(pga, oracle):
    reqs = _requests(oracle, 300, [20, 75])
    server = pga.Server(64 * MiB, "batched", chol_uplo="U")
    got, _ = _run(server, reqs)
    server.close()
    for (kind, n, A, B), g in zip(reqs, got):
        F = A.copy(order="F")
        X = B.copy(order="F")
        if kind == "lu":
            _, piv, info = oracle.rgetrf(F, nb=64, nthreads=THREADS)
            oracle.rgetrs("N", F, piv, X)
        else:
            _, info = oracle.rpotrf("U", F, nb=64, nthreads=THREADS)
            oracle.rpotrs("U", F, X)
        assert g["info"] == info == 0
        assert np.array_equal(g["x"], X)


def test_singular_and_not_pd(pga, oracle):
    n = 90
    Z = posit_matrix(oracle, 400, n, n)
    Z[:, 17] = 0
    P = _spd(oracle, 401, n)
    P[50, 50] = oracle.from_double(-3.0)
    B = posit_matrix(oracle, 402, n, 1)
    reqs = [("lu", n, Z, B), ("chol", n, P, B), ("lu", 700, posit_matrix(oracle, 403, 700, 700), None)]
    reqs[2] = ("lu", 700, reqs[2][2], posit_matrix(oracle, 404, 700, 1))
    reqs[2][2][:, 5] = 0
    for mode in ("batched", "plain"):
        s = pga.Server(64 * MiB, mode, chol_uplo="U")
        got, _ = _run(s, reqs)
        s.close()
        assert got[0]["info"] == 18 and got[1]["info"] == 51 and got[2]["info"] == 6
        assert all(g["status"] == pga.REQ_OK for g in got)


def test_budget_admission(pga, oracle):
    """A budget that holds only a few requests at once still serves them all, never
    above the budget; a request that can never fit is refused."""
    budget = 2 * MiB
    reqs = _requests(oracle, 500, [400] * 10)  # ~0.65 MiB each: three fit at a time
    s = pga.Server(budget, "batched", batch_max_n=512)
    got, _ = _run(s, reqs)
    rid = s.submit("lu", 2000, posit_matrix(oracle, 510, 2000, 2000), posit_matrix(oracle, 511, 2000, 1))
    big = s.wait(rid)
    st = s.stats()
    s.close()
    assert all(g["status"] == pga.REQ_OK for g in got)
    assert st["bytes_peak"] <= budget
    assert big["status"] == pga.REQ_TOO_LARGE
    plain = pga.Server(64 * MiB, "plain")
    want, _ = _run(plain, reqs)
    plain.close()
    for g, w in zip(got, want):
        assert np.array_equal(g["x"], w["x"])


def test_resident_factors_and_eviction(pga, oracle):
    n = 300  # factor ~0.35 MiB
    budget = 2 * MiB
    s = pga.Server(budget, "batched", batch_max_n=512)
    mats = [posit_matrix(oracle, 600 + i, n, n) for i in range(6)]
    B = posit_matrix(oracle, 620, n, 2)
    ids = []
    for A in mats:  # one at a time: each keeps its factor; later ones evict the oldest
        ids.append(s.submit("lu", n, A, B))
        assert s.wait(ids[-1])["status"] == pga.REQ_OK
    st = s.stats()
    assert st["evictions"] >= 1
    first = s.wait(s.submit("resolve", n, b=B, factor=ids[0]))
    last = s.wait(s.submit("resolve", n, b=B, factor=ids[-1]))
    s.close()
    assert first["status"] == pga.REQ_NOT_RESIDENT
    assert last["status"] == pga.REQ_OK
    plain = pga.Server(64 * MiB, "plain")
    want = plain.wait(plain.submit("lu", n, mats[-1], B))
    plain.close()
    assert np.array_equal(last["x"], want["x"])


def test_resolve_waits_for_its_factor(pga, oracle):
    n = 800  # worker path: the factorization is still running when the re-solve arrives
    A = posit_matrix(oracle, 700, n, n)
    B1, B2 = posit_matrix(oracle, 701, n, 1), posit_matrix(oracle, 702, n, 3)
    s = pga.Server(256 * MiB, "batched")
    i1 = s.submit("lu", n, A, B1)
    i2 = s.submit("resolve", n, b=B2, factor=i1)
    r1, r2 = s.wait(i1), s.wait(i2)
    s.close()
    assert r1["status"] == r2["status"] == pga.REQ_OK
    assert r2["t_admit"] >= r1["t_done"] - 1e-3 or r2["t_done"] >= r1["t_done"]
    plain = pga.Server(256 * MiB, "plain")
    w1 = plain.submit("lu", n, A, B2)
    want = plain.wait(w1)
    plain.close()
    assert np.array_equal(r2["x"], want["x"])


def test_failed_launches_finish_as_failed(pga, oracle, monkeypatch):
    """Launch failures (forced from the host before anything is enqueued, every second
    launch of a batch slot or worker) finish their requests as failed without touching
    the others: the rest stay bit-exact, a failed factor is never kept for re-solves, and
    the server keeps serving afterwards."""
    monkeypatch.setenv("PGA_SERVE_FAIL_EVERY", "2")
    sizes = [30, 600, 120, 530, 64, 700]  # batch slots and workers alternate
    reqs = _requests(oracle, 800, sizes)
    s = pga.Server(256 * MiB, "batched", batch_max_n=512)
    monkeypatch.delenv("PGA_SERVE_FAIL_EVERY")
    ids, got = [], []
    for kind, n, A, B in reqs:  # one at a time: one launch per request
        ids.append(s.submit(kind, n, A, B))
        got.append(s.wait(ids[-1]))
    gone = s.wait(s.submit("resolve", sizes[1], b=reqs[1][3], factor=ids[1]))
    after = s.wait(s.submit(*reqs[0]))  # the seventh launch is not forced to fail
    s.close()
    plain = pga.Server(256 * MiB, "plain")
    want = [plain.wait(plain.submit(kind, n, A, B)) for kind, n, A, B in reqs]
    plain.close()
    for i, (g, w) in enumerate(zip(got, want)):
        if i % 2:
            assert g["status"] == pga.REQ_FAILED, i
        else:
            assert g["status"] == w["status"] == pga.REQ_OK, i
            assert np.array_equal(g["x"], w["x"]), i
    assert gone["status"] == pga.REQ_NOT_RESIDENT
    assert after["status"] == pga.REQ_OK and np.array_equal(after["x"], want[0]["x"])
