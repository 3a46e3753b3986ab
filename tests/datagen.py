"""Seeded synthetic inputs, restating the device generator of the product library
(pga_fill_uniform_p32_device): u_i = SplitMix64(seed, i) >> 11 scaled by 2^-53,
x_i = lo + (hi - lo) * u_i, then posit32 round-to-nearest-even (done by the oracle)."""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def uniform01(seed, n, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniform_doubles(seed, n, lo=-1.0, hi=1.0):
    return lo + (hi - lo) * uniform01(seed, n)


def posit_matrix(oracle, seed, rows, cols, lo=-1.0, hi=1.0):
    """Column-major uint32 matrix of posit32(uniform[lo, hi)) from the seed."""
    x = uniform_doubles(seed, rows * cols, lo, hi)
    return np.asfortranarray(oracle.from_double_v(x).reshape(cols, rows).T)


def random_patterns(rng, n, long_regimes=False):
    a = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    if long_regimes:
        s = rng.integers(0, 31, n).astype(np.uint32)
        a = (a >> s) | (a & np.uint32(0x80000000))
    return a
