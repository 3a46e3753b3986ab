"""Write tests/golden/p32_addmul_ref.npz: input pairs and the outputs of the reference's
own p32_add / p32_mul (libRpack/kernel_entry.cl:370, :399), built for the host by
`make -C oracle ref` into oracle/_ref/libposit_ref.so.

Run in a container that has the reference tree:  python tests/golden/gen_posit_golden.py
The pairs cover: every product of a set of edge patterns (0, NaR, +-1, +-minpos,
+-maxpos and their neighbours, regime-boundary powers of two, values with regime
lengths 26..30), results that land on regime 29/30, near-cancellation pairs and
uniform random patterns.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import RefPosit  # noqa: E402


def edge_patterns():
    s = set()
    for v in (0, 0x80000000, 0x40000000, 0x7FFFFFFF, 0x7FFFFFFE, 0x1, 0x2, 0x3, 0x40000001, 0x3FFFFFFF,
              0x48000000, 0x38000000, 0x50000000, 0x30000000, 0x7FFF0000, 0x00010000, 0x7FFFFFFC, 0x4):
        s.add(v)
        s.add((-v) & 0xFFFFFFFF)
    # powers of two 2^S for S in [-120, 120]: pattern for regime k, exponent e, fraction 0
    for S in range(-120, 121, 1):
        k, e = divmod(S, 4)
        if k >= 0:
            r = k + 1
            bits = "1" * r + "0" + format(e, "02b")
        else:
            r = -k
            bits = "0" * r + "1" + format(e, "02b")
        bits = (bits + "0" * 31)[:31]
        p = int(bits, 2)
        if S % 7 == 0 or abs(S) > 100 or abs(S) < 8:
            s.add(p)
            s.add((-p) & 0xFFFFFFFF)
            s.add((p + 1) & 0xFFFFFFFF)
            s.add((p - 1) & 0xFFFFFFFF)
    return np.array(sorted(s), dtype=np.uint32)


def main():
    R = RefPosit()
    rng = np.random.default_rng(20240131)
    e = edge_patterns()
    ea = np.repeat(e, e.size)
    eb = np.tile(e, e.size)
    pick = rng.choice(ea.size, size=min(ea.size, 6000), replace=False)
    ea, eb = ea[pick], eb[pick]
    # uniform random patterns and long-regime patterns
    ra = rng.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    rb = rng.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    sh = rng.integers(0, 31, 3000).astype(np.uint32)
    la = (ra >> sh) | (ra & 0x80000000)
    lb = (rb >> rng.integers(0, 31, 3000).astype(np.uint32)) | (rb & 0x80000000)
    # near cancellation: b close to -a
    ca = rng.integers(0, 2**32, 2000, dtype=np.uint64).astype(np.uint32)
    cb = ((-(ca.astype(np.int64)) + rng.integers(-64, 64, 2000)) & 0xFFFFFFFF).astype(np.uint32)
    a = np.concatenate([ea, ra, la, ca])
    b = np.concatenate([eb, rb, lb, cb])
    add = R.add_v(a, b)
    mul = R.mul_v(a, b)
    out = os.path.join(HERE, "p32_addmul_ref.npz")
    np.savez_compressed(out, a=a, b=b, add=add, mul=mul)
    print("wrote", out, a.size, "pairs")


if __name__ == "__main__":
    main()
