"""Every operator of the drop-in posit32 C++ type (include/softposit_cpp.h) against the
oracle, through a C++ program that uses the type the way the reference code does
(tests/cpp/posit_ops_driver.cpp).  The type's arithmetic runs on the host through the
library's scalar entry points, so this needs no GPU."""
import os
import subprocess

import numpy as np
import pytest

from datagen import random_patterns

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "build", "posit_ops_driver")
NAR = 0x80000000
NOPS = 21
IN_T = np.dtype([("a", "<u4"), ("b", "<u4"), ("d", "<f8")])
OUT_T = np.dtype([("r", "<u4", (NOPS,)), ("todouble_a", "<f8")], align=True)  # C struct layout


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "all"], check=True, capture_output=True)
    return DRIVER


def _records(n, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros(n, IN_T)
    q = n // 4
    rec["a"][:q] = random_patterns(rng, q)
    rec["a"][q:2 * q] = random_patterns(rng, q, long_regimes=True)
    rec["a"][2 * q:] = random_patterns(rng, n - 2 * q)
    rec["b"] = random_patterns(rng, n)
    rec["b"][q:2 * q] = random_patterns(rng, q, long_regimes=True)
    special = np.array([0, NAR, 0x40000000, 0xC0000000, 0x7FFFFFFF, 0x00000001, 0x80000001, 0xFFFFFFFF],
                       np.uint32)
    k = special.size
    rec["a"][:k * k] = np.repeat(special, k)
    rec["b"][:k * k] = np.tile(special, k)
    d = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-130, 130, n)
    d[: n // 8] = rng.uniform(-3, 3, n // 8)
    d[n // 8: n // 8 + 6] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-300]
    rec["d"] = d
    return rec


def test_every_operator_matches_oracle(driver, oracle, tmp_path):
    n = 200_000
    rec = _records(n, 11)
    rec.tofile(tmp_path / "in.bin")
    subprocess.run([driver, str(tmp_path / "in.bin"), str(n), str(tmp_path / "out.bin")], check=True, timeout=120)
    out = np.fromfile(tmp_path / "out.bin", OUT_T)
    assert out.size == n
    a, b, d = rec["a"], rec["b"], rec["d"]
    O = oracle
    pd = O.from_double_v(d)
    fin = np.isfinite(d) & (np.abs(d) < 1e6)
    di1000 = np.where(fin, np.trunc(np.where(fin, d, 0.0) * 1000.0), 0.0)
    di7 = np.where(fin, np.trunc(np.where(fin, d, 0.0) * 7.0), 0.0)
    with np.errstate(over="ignore"):
        df = d.astype(np.float32).astype(np.float64)
    sa, sb = a.view(np.int32), b.view(np.int32)
    cmp = ((a == b).astype(np.uint32) | (a != b) << 1 | (sa < sb) << 2 | (sa <= sb) << 3 | (sa > sb) << 4
           | (sa >= sb) << 5 | (a == NAR) << 6).astype(np.uint32)
    want = [
        ("a + b", O.add_v(a, b)), ("a - b", O.sub_v(a, b)), ("a * b", O.mul_v(a, b)), ("a / b", O.div_v(a, b)),
        ("-a", ((-a.astype(np.int64)) & 0xFFFFFFFF).astype(np.uint32)),
        ("x += b", O.add_v(a, b)), ("x -= b", O.sub_v(a, b)), ("x *= b", O.mul_v(a, b)), ("x /= b", O.div_v(a, b)),
        ("sqrt(a)", O.sqrt_v(a)), ("x.sqrt()", O.sqrt_v(a)),
        ("posit32(double)", pd), ("x = double", pd),
        ("double + b", O.add_v(pd, b)), ("double - b", O.sub_v(pd, b)), ("double * b", O.mul_v(pd, b)),
        ("double / b", O.div_v(pd, b)),
        ("posit32(int)", O.from_double_v(di1000)), ("x = int", O.from_double_v(di7)),
        ("posit32(float)", O.from_double_v(df)),
        ("== != < <= > >= isNaR", cmp),
    ]
    assert len(want) == NOPS
    for col, (name, w) in enumerate(want):
        bad = np.flatnonzero(out["r"][:, col] != w)
        assert bad.size == 0, f"{name}: {bad.size} mismatches, first a={a[bad[0]]:#x} b={b[bad[0]]:#x} d={d[bad[0]]!r}"
    td = O.to_double_v(a)
    same = (out["todouble_a"] == td) | (np.isnan(out["todouble_a"]) & np.isnan(td))
    assert same.all(), "toDouble"
    assert np.isnan(out["todouble_a"][a == NAR]).all()
