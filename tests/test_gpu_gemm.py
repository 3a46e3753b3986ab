"""GPU parity of the posit32 GEMM (through the C ABI) against the oracle's restatement of
the reference CPU Rgemm (libRpack/Rgemm.cpp:62-251).  Bit-exact: every output pattern
must match."""
import os

import numpy as np
import pytest

from datagen import posit_matrix, random_patterns

pytestmark = pytest.mark.gpu

ONE, MONE, NAR = 0x40000000, 0xC0000000, 0x80000000
THREADS = min(16, os.cpu_count() or 1)


def _both(oracle, pga, ta, tb, alpha, beta, A, B, C, impl="tiled"):
    m, n = C.shape
    k = A.shape[0] if ta == "T" else A.shape[1]
    want = C.copy(order="F")
    assert oracle.rgemm(ta, tb, m, n, k, alpha, A, B, beta, want, nthreads=THREADS) == 0
    pga.set_gemm_impl(impl)
    try:
        got = pga.rgemm(ta, tb, alpha, A, B, beta, C)
    finally:
        pga.set_gemm_impl("tiled")
    return got, want


def _assert_same(got, want):
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first {[(tuple(i), hex(got[tuple(i)]), hex(want[tuple(i)])) for i in bad[:4]]}"


def test_config1_512_nn(pga, oracle):
    """Plain Rgemm C = A*B, M=N=K=512, alpha=1, beta=0, uniform[-1,1) inputs."""
    n = 512
    A = posit_matrix(oracle, 101, n, n)
    B = posit_matrix(oracle, 102, n, n)
    C = posit_matrix(oracle, 103, n, n)
    got, want = _both(oracle, pga, "N", "N", ONE, 0, A, B, C)
    _assert_same(got, want)


@pytest.mark.parametrize("impl", ["tiled", "exact"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("alpha,beta", [(ONE, 0), (MONE, ONE), (0x3A000000, 0xC8000000), (ONE, ONE), (0, 0x44000000)])
def test_variants_ragged(pga, oracle, impl, ta, tb, alpha, beta):
    m, n, k = 77, 45, 133
    A = posit_matrix(oracle, 1, k if ta == "T" else m, m if ta == "T" else k)
    B = posit_matrix(oracle, 2, n if tb == "T" else k, k if tb == "T" else n)
    C = posit_matrix(oracle, 3, m, n)
    got, want = _both(oracle, pga, ta, tb, alpha, beta, A, B, C, impl)
    _assert_same(got, want)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (31, 200, 17), (200, 31, 16), (64, 64, 0), (64, 64, 1), (65, 129, 257),
                                   (3, 300, 5), (130, 2, 64)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
def test_shapes(pga, oracle, m, n, k, ta, tb):
    # operands stored with at least one row/column so that k == 0 still has valid buffers
    A = posit_matrix(oracle, 11, max(1, k if ta == "T" else m), max(1, m if ta == "T" else k))
    B = posit_matrix(oracle, 12, max(1, n if tb == "T" else k), max(1, k if tb == "T" else n))
    C = posit_matrix(oracle, 13, m, n)
    for alpha, beta in [(MONE, ONE), (0x3C000000, 0x38000000)]:
        want = C.copy(order="F")
        got = C.copy(order="F")
        assert oracle.rgemm(ta, tb, m, n, k, alpha, A, B, beta, want, nthreads=THREADS) == 0
        pga.RgemmG(ta, tb, m, n, k, alpha, A, max(1, A.shape[0]), B, max(1, B.shape[0]), beta, got, max(1, m))
        _assert_same(got, want)


@pytest.mark.parametrize("kind", ["patterns", "long_regimes", "sparse", "nar", "tiny", "huge", "cancel"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_value_classes(pga, oracle, kind, ta, tb):
    """Inputs that push the decoded fast path to its limits or out of it."""
    rng = np.random.default_rng(hash(kind) % 1000 + ord(ta) + 3 * ord(tb))
    m, n, k = 70, 66, 96

    def mat(r, c, seed):
        if kind == "patterns":
            v = random_patterns(rng, r * c)
        elif kind == "long_regimes":
            v = random_patterns(rng, r * c, True)
        elif kind == "sparse":
            v = posit_matrix(oracle, seed, r, c).ravel(order="F").copy()
            v[rng.random(v.size) < 0.7] = 0
        elif kind == "nar":
            v = posit_matrix(oracle, seed, r, c).ravel(order="F").copy()
            v[rng.integers(0, v.size, 3)] = NAR
        elif kind == "tiny":  # products near 2^-100, sums can cancel below 2^-108
            v = oracle.from_double_v(rng.uniform(-1, 1, r * c) * 2.0 ** -50)
        elif kind == "huge":  # products near 2^100: accumulators leave the fast range
            v = oracle.from_double_v(rng.uniform(-1, 1, r * c) * 2.0 ** 52)
        else:  # cancel: few distinct values so that many sums cancel exactly
            v = oracle.from_double_v(rng.choice([-1.0, -0.5, 0.5, 1.0, 0.25], r * c))
        return np.asfortranarray(np.asarray(v, np.uint32).reshape(c, r).T)

    A = mat(k if ta == "T" else m, m if ta == "T" else k, 1)
    B = mat(n if tb == "T" else k, k if tb == "T" else n, 2)
    C = mat(m, n, 3)
    for alpha, beta in [(ONE, 0), (MONE, ONE)]:
        got, want = _both(oracle, pga, ta, tb, alpha, beta, A, B, C)
        _assert_same(got, want)


def test_tiled_equals_exact_kernel(pga, oracle):
    rng = np.random.default_rng(77)
    m, n, k = 300, 260, 200
    A = posit_matrix(oracle, 31, m, k, -4, 4)
    B = posit_matrix(oracle, 32, k, n, -4, 4)
    C = posit_matrix(oracle, 33, m, n, -4, 4)
    pga.set_gemm_impl("exact")
    e = pga.rgemm("N", "N", MONE, A, B, ONE, C)
    pga.set_gemm_impl("tiled")
    t = pga.rgemm("N", "N", MONE, A, B, ONE, C)
    _assert_same(t, e)
    del rng


def test_device_submatrix_update(pga, oracle):
    """Trailing update on a sub-block of a resident matrix: C22 -= A21 * A12, ld = n."""
    n, j, jb = 700, 128, 64
    M = posit_matrix(oracle, 41, n, n)
    d = pga.DeviceArray.from_host(M)
    r0 = j + jb
    mm = nn = n - r0
    pga.rgemm_device("N", "N", mm, nn, jb, MONE, d.offset(r0 + j * n), n, d.offset(j + r0 * n), n, ONE,
                     d.offset(r0 + r0 * n), n)
    pga.synchronize()
    got = d.to_host()
    want = M.copy(order="F")
    sub = np.asfortranarray(want[r0:, r0:])
    oracle.rgemm("N", "N", mm, nn, jb, MONE, np.asfortranarray(want[r0:, j:j + jb]),
                 np.asfortranarray(want[j:j + jb, r0:]), ONE, sub, nthreads=THREADS)
    want[r0:, r0:] = sub
    _assert_same(got, want)


def test_illegal_arguments(pga):
    A = np.zeros((4, 4), np.uint32, order="F")
    with pytest.raises(pga.PositError) as e:
        pga.RgemmG("X", "N", 4, 4, 4, ONE, A, 4, A, 4, 0, A.copy(order="F"), 4)
    assert e.value.argument == 1
    with pytest.raises(pga.PositError) as e:
        pga.RgemmG("N", "N", 4, 4, 4, ONE, A, 3, A, 4, 0, A.copy(order="F"), 4)
    assert e.value.argument == 8


def test_trailing_shape_4096_full(pga, oracle):
    """Config-2 shape M=N=4096, K=256, alpha=-1, beta=1, device-resident, whole result."""
    M = N = 4096
    K = 256
    dA = pga.DeviceArray((M, K))
    dB = pga.DeviceArray((K, N))
    dC = pga.DeviceArray((M, N))
    pga.fill_uniform(dA, 201)
    pga.fill_uniform(dB, 202)
    pga.fill_uniform(dC, 203)
    A, B, C = dA.to_host(), dB.to_host(), dC.to_host()
    pga.rgemm_device("N", "N", M, N, K, MONE, dA, M, dB, K, ONE, dC, M)
    pga.synchronize()
    got = dC.to_host()
    oracle.rgemm("N", "N", M, N, K, MONE, A, B, ONE, C, nthreads=THREADS)
    _assert_same(got, C)


def test_trailing_shape_16384_sampled_columns(pga, oracle):
    """Config-2 shape M=N=16384, K=256: full-size launch, every row of 48 random columns
    checked against the oracle (each column of C depends only on its own column of B)."""
    M = N = 16384
    K = 256
    dA = pga.DeviceArray((M, K))
    dB = pga.DeviceArray((K, N))
    dC = pga.DeviceArray((M, N))
    pga.fill_uniform(dA, 301)
    pga.fill_uniform(dB, 302)
    pga.fill_uniform(dC, 303)
    A, B, C0 = dA.to_host(), dB.to_host(), dC.to_host()
    pga.rgemm_device("N", "N", M, N, K, MONE, dA, M, dB, K, ONE, dC, M)
    pga.synchronize()
    got = dC.to_host()
    cols = np.sort(np.random.default_rng(9).choice(N, 48, replace=False))
    Bs = np.asfortranarray(B[:, cols])
    Cs = np.asfortranarray(C0[:, cols])
    oracle.rgemm("N", "N", M, cols.size, K, MONE, A, Bs, ONE, Cs, nthreads=THREADS)
    _assert_same(np.asfortranarray(got[:, cols]), Cs)


def test_tile_override_same_bits(pga, tmp_path):
    """PGA_GEMM_TILE fixes the plain GEMM's tile edge (the role of OPENCL_GEMM_BLOCKSIZE);
    a process started with each value computes the same bits as the per-shape choice."""
    import os
    import subprocess
    import sys
    m, n, k = 300, 270, 130
    script = (
        "import sys, numpy as np; sys.path.insert(0, {root!r}); import posit_gpu_amd as pga; pga.init(0)\n"
        "A = pga.DeviceArray(({m}, {k})); B = pga.DeviceArray(({k}, {n})); C = pga.DeviceArray(({m}, {n}))\n"
        "pga.fill_uniform(A, 91); pga.fill_uniform(B, 92); pga.fill_uniform(C, 93)\n"
        "pga.rgemm_device('N', 'N', {m}, {n}, {k}, 0xC0000000, A, {m}, B, {k}, 0x40000000, C, {m})\n"
        "pga.synchronize(); np.save(sys.argv[1], C.to_host())\n"
    ).format(root=os.path.dirname(os.path.dirname(os.path.abspath(__file__))), m=m, n=n, k=k)
    outs = {}
    for tile in ("", "16", "32", "64"):
        env = dict(os.environ)
        env.pop("PGA_GEMM_TILE", None)
        if tile:
            env["PGA_GEMM_TILE"] = tile
        out = tmp_path / f"c{tile or 'auto'}.npy"
        subprocess.run([sys.executable, "-c", script, str(out)], env=env, check=True, timeout=240)
        outs[tile] = np.load(out)
    for tile in ("16", "32", "64"):
        assert np.array_equal(outs[tile], outs[""]), tile


@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 300, 270, 130), ("T", "N", 64, 1000, 33), ("N", "T", 777, 5, 260)])
def test_autotuned_tile_same_bits(pga, oracle, ta, tb, m, n, k):
    """pga_rgemm_autotune times the tile edges on one shape and keeps the fastest for it;
    calls of that shape then use it and give the oracle's bits."""
    rng = np.random.default_rng(m + n + k)
    A = posit_matrix(oracle, 141, *((k, m) if ta == "T" else (m, k)))
    B = posit_matrix(oracle, 142, *((n, k) if tb == "T" else (k, n)))
    C = posit_matrix(oracle, 143, m, n)
    C[rng.integers(0, m, 3), rng.integers(0, n, 3)] = NAR
    before = pga.rgemm(ta, tb, MONE, A, B, ONE, C)
    tile = pga.rgemm_autotune(ta, tb, m, n, k)
    assert tile in (16, 32, 64)
    got, want = _both(oracle, pga, ta, tb, MONE, ONE, A, B, C)
    _assert_same(got, want)
    _assert_same(before, want)


_TILE_SCRIPT = """
import sys, numpy as np
sys.path.insert(0, {root!r})
import posit_gpu_amd as pga
pga.init(0)
d = np.load(sys.argv[1])
out = {{}}
for key in d.files:
    if not key.startswith('C_'):
        continue
    case = key[2:]
    kind, ta, tb, alpha, beta = case.split('-')
    A, B, C = d['A_' + kind + ta], d['B_' + kind + tb], d[key]
    out[case] = pga.rgemm(ta, tb, int(alpha), np.asfortranarray(A), np.asfortranarray(B), int(beta),
                          np.asfortranarray(C))
np.savez(sys.argv[2], **out)
"""


def _class_matrix(oracle, rng, kind, r, c, seed):
    if kind == "patterns":
        v = random_patterns(rng, r * c)
    elif kind == "long_regimes":
        v = random_patterns(rng, r * c, True)
    elif kind == "sparse":
        v = posit_matrix(oracle, seed, r, c).ravel(order="F").copy()
        v[rng.random(v.size) < 0.7] = 0
    elif kind == "nar":
        v = posit_matrix(oracle, seed, r, c).ravel(order="F").copy()
        v[rng.integers(0, v.size, 3)] = NAR
    elif kind == "tiny":  # products near 2^-100, sums can cancel below 2^-108
        v = oracle.from_double_v(rng.uniform(-1, 1, r * c) * 2.0 ** -50)
    elif kind == "huge":  # products near 2^100: accumulators leave the fast range
        v = oracle.from_double_v(rng.uniform(-1, 1, r * c) * 2.0 ** 52)
    elif kind == "boundary":  # products 2^106..2^108: running sums cross 2^108 inside a slice
        v = oracle.from_double_v(rng.choice([-1.0, 1.0], r * c) * rng.uniform(0.5, 1.0, r * c) * 2.0 ** 54)
    else:  # cancel: few distinct values so that many sums cancel exactly
        v = oracle.from_double_v(rng.choice([-1.0, -0.5, 0.5, 1.0, 0.25], r * c))
    return np.asfortranarray(np.asarray(v, np.uint32).reshape(c, r).T)


@pytest.mark.parametrize("tile", ["64", "32", "16", "64-panels"])
def test_value_classes_at_each_tile(pga, oracle, tmp_path, tile):
    """Every value class through the tile that serves large GEMMs: a process started with
    PGA_GEMM_TILE fixed runs all four variants and three (alpha, beta) pairs, with K long
    enough (288) that the one-element tile also takes the FP64 form.  This exercises the
    FP64 form's slice redo (NaN from a scale outside the table, NaR) and the sum-first
    epilogue of TN / TT at 4 x 2, 2 x 2 and 1 x 1 elements per lane, against the oracle."""
    import subprocess
    import sys
    m, n, k = 70, 66, 288
    kinds = ["patterns", "long_regimes", "sparse", "nar", "tiny", "huge", "boundary", "cancel"]
    pairs = [(ONE, 0), (MONE, ONE), (0x3A000000, 0xC8000000)]
    arrays, want = {}, {}
    for ci, kind in enumerate(kinds):
        rng = np.random.default_rng(500 + ci)
        A = _class_matrix(oracle, rng, kind, m, k, 1)
        B = _class_matrix(oracle, rng, kind, k, n, 2)
        C = _class_matrix(oracle, rng, kind, m, n, 3)
        arrays[f"A_{kind}N"], arrays[f"A_{kind}T"] = A, np.asfortranarray(A.T)
        arrays[f"B_{kind}N"], arrays[f"B_{kind}T"] = B, np.asfortranarray(B.T)
        for ta in "NT":
            for tb in "NT":
                for alpha, beta in pairs:
                    case = f"{kind}-{ta}-{tb}-{alpha}-{beta}"
                    arrays["C_" + case] = C
                    w = C.copy(order="F")
                    assert oracle.rgemm(ta, tb, m, n, k, alpha, arrays[f"A_{kind}{ta}"], arrays[f"B_{kind}{tb}"],
                                        beta, w, nthreads=THREADS) == 0
                    want[case] = w
    np.savez(tmp_path / "in.npz", **arrays)
    env = dict(os.environ, PGA_GEMM_TILE=tile.split("-")[0])
    if tile.endswith("panels"):  # the 64 x 32 tile with its operands from FP64 panels
        env["PGA_GEMM_PANEL_MIN"] = "0"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([sys.executable, "-c", _TILE_SCRIPT.format(root=root), str(tmp_path / "in.npz"),
                    str(tmp_path / "out.npz")], env=env, check=True, timeout=240)
    got = np.load(tmp_path / "out.npz")
    assert sorted(got.files) == sorted(want)
    for case, w in want.items():
        bad = np.argwhere(got[case] != w)
        assert bad.size == 0, f"tile {tile} {case}: {len(bad)} mismatches, first {bad[:3].tolist()}"


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_panels_ragged_vs_oracle(pga, oracle, ta, tb):
    """A GEMM above the panel threshold (m*n*k >= 2^30) with ragged edges in m, n and K:
    the 64 x 32 tile takes its operands from FP64 panels; every element against the oracle."""
    m, n, k = 1500, 1500, 500
    A = posit_matrix(oracle, 61, k if ta == "T" else m, m if ta == "T" else k)
    B = posit_matrix(oracle, 62, n if tb == "T" else k, k if tb == "T" else n)
    C = posit_matrix(oracle, 63, m, n)
    for alpha, beta in [(MONE, ONE), (0x3A000000, 0xC8000000)]:
        got, want = _both(oracle, pga, ta, tb, alpha, beta, A, B, C)
        _assert_same(got, want)
