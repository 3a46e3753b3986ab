"""GPU parity of the scalar posit32 core (pattern-level operations and codecs) against
the oracle, through the C ABI."""
import numpy as np
import pytest

from datagen import random_patterns, uniform_doubles

pytestmark = pytest.mark.gpu


def _patterns(rng, n):
    edge = np.array([0, 0x80000000, 0x40000000, 0xC0000000, 0x7FFFFFFF, 0x80000001, 1, 0xFFFFFFFF,
                     2, 0x7FFFFFFE, 0x40000001, 0x3FFFFFFF], np.uint32)
    return np.concatenate([edge, random_patterns(rng, n // 2), random_patterns(rng, n // 2, True)])


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_binary_ops(pga, oracle, op):
    rng = np.random.default_rng({"add": 1, "sub": 2, "mul": 3, "div": 4}[op])
    a = _patterns(rng, 1 << 20)
    b = _patterns(rng, 1 << 20)
    got = pga.elementwise(op, a, b)
    want = getattr(oracle, op + "_v")(a, b)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(hex(a[i]), hex(b[i]), hex(got[i]), hex(want[i])) for i in bad[:5]]


def test_golden_vectors_on_gpu(pga):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "p32_addmul_ref.npz"))
    assert np.array_equal(pga.elementwise("add", g["a"], g["b"]), g["add"])
    assert np.array_equal(pga.elementwise("mul", g["a"], g["b"]), g["mul"])


def test_sqrt(pga, oracle):
    rng = np.random.default_rng(5)
    a = _patterns(rng, 1 << 19)
    assert np.array_equal(pga.elementwise("sqrt", a), oracle.sqrt_v(a))


def test_codecs(pga, oracle):
    rng = np.random.default_rng(6)
    d = np.concatenate([uniform_doubles(7, 1 << 18), rng.standard_normal(1 << 18) * 10.0 ** rng.integers(-45, 45, 1 << 18),
                        [0.0, -0.0, 1e300, -1e300, 5e-324, float("inf"), float("nan"), 2.0 ** 119 * 1.5]])
    p = pga.from_double(d)
    assert np.array_equal(p, oracle.from_double_v(d))
    q = _patterns(rng, 1 << 18)
    x = pga.to_double(q)
    y = oracle.to_double_v(q)
    assert np.array_equal(np.isnan(x), np.isnan(y))
    assert np.array_equal(x[~np.isnan(x)], y[~np.isnan(y)])


def test_device_generator_matches_host_restatement(pga, oracle):
    n = 100_003
    d = pga.DeviceArray(n, np.uint32, "C")
    pga.fill_uniform(d, seed=1234)
    pga.synchronize()
    assert np.array_equal(d.to_host(), oracle.from_double_v(uniform_doubles(1234, n)))
