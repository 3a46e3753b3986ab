"""Writes tests/golden/fullsize_hashes.json: SHA-256 digests of the oracle's results at the
full sizes of the configurations, for the GPU test that compares the device results with
them (tests/test_gpu_fullsize.py).

Run in the build container (it takes the better part of an hour on 8 cores):

    python tests/golden/gen_fullsize_hashes.py [--threads 8] [--only gemm|lu|chol]

Workloads (inputs are reproduced on the device by pga_fill_uniform_p32_device with the same
seeds; datagen.py restates that generator):
  gemm  C = C - A*B (Rgemm 'N','N', alpha = -1, beta = 1) for M = N in {16384, 12288, 8192,
        4096}, K = 256: the trailing-update shapes of a blocked LU at n = 16384, nb = 256.
        A (M x K), B (K x N), C (M x N) uniform[-1,1) with seeds 1000+3i, 1001+3i, 1002+3i.
  lu    Rgetrf (nb = 256) of A = uniform[-1,1) n = 8192, seed 4242; b = A*1 by Rgemv 'N'
        (alpha = 1, beta = 0); Rgetrs 'N' of b.
  chol  Rpotrf 'L' (nb = 64, iMlaenv's) of S = M*M^T/n + I, M uniform[-1,1) (seed 4300)
        truncated to multiples of 2^-20 so that M*M^T is exact in FP64 whatever the
        summation order (each product is an integer number of units of 2^-40, at most
        2^40 of them, so every partial sum of 8192 products stays within 2^53 units);
        S / n and + I are single IEEE operations; then posit32 RNE.  b = S*1 by Rgemv 'N';
        Rpotrs 'L' of b.

Every matrix is hashed as its column-major uint32 little-endian bytes; besides the whole
array each 1024-column block is hashed so that a mismatch can be located.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))  # tests/

from datagen import posit_matrix, uniform01  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

OUT = os.path.join(HERE, "fullsize_hashes.json")
ONE, MONE = 0x40000000, 0xC0000000
GEMM_SHAPES = [16384, 12288, 8192, 4096]
GEMM_K = 256
LU_N, LU_NB, LU_SEED = 8192, 256, 4242
CHOL_N, CHOL_NB, CHOL_SEED = 8192, 64, 4300
BLOCK = 1024


def digest(a):
    a = np.asarray(a)
    if a.ndim == 2:
        b = np.asfortranarray(a).tobytes(order="F")
    else:
        b = np.ascontiguousarray(a).tobytes()
    return hashlib.sha256(b).hexdigest()


def digests(a):
    """Whole-array digest plus one per BLOCK columns (2-D) or BLOCK entries (1-D)."""
    a = np.asarray(a)
    if a.ndim == 2:
        parts = [digest(a[:, j:j + BLOCK]) for j in range(0, a.shape[1], BLOCK)]
    else:
        parts = [digest(a[j:j + BLOCK]) for j in range(0, a.shape[0], BLOCK)]
    return {"sha256": digest(a), "blocks": parts, "block": BLOCK}


def spd_quantized(oracle, seed, n):
    """S = M*M^T/n + I in FP64 with M uniform[-1,1) truncated to multiples of 2^-20, then
    posit32.  M*M^T is exact in FP64 for any summation order (integers below 2^53 in units
    of 2^-40), so every machine builds the same S."""
    u = uniform01(seed, n * n)
    M = (np.floor(u * 2.0 ** 21) - 2.0 ** 20).reshape(n, n)  # integers in [-2^20, 2^20)
    S = M @ M.T  # exact integers (units of 2^-40)
    S *= 2.0 ** -40
    S /= n
    S[np.diag_indices(n)] += 1.0
    return np.asfortranarray(oracle.from_double_v(S.ravel(order="F")).reshape(n, n, order="F"))


def run_gemm(O, threads, res):
    for i, n in enumerate(GEMM_SHAPES):
        t0 = time.time()
        A = posit_matrix(O, 1000 + 3 * i, n, GEMM_K)
        B = posit_matrix(O, 1001 + 3 * i, GEMM_K, n)
        C = posit_matrix(O, 1002 + 3 * i, n, n)
        ent = {"m": n, "n": n, "k": GEMM_K, "seeds": [1000 + 3 * i, 1001 + 3 * i, 1002 + 3 * i],
               "a": digest(A), "b": digest(B), "c_in": digest(C)}
        assert O.rgemm("N", "N", n, n, GEMM_K, MONE, A, B, ONE, C, nthreads=threads) == 0
        ent["c_out"] = digests(C)
        ent["oracle_s"] = round(time.time() - t0, 1)
        res["gemm"][str(n)] = ent
        print(f"gemm {n}: {ent['oracle_s']} s", flush=True)
        save(res)


def run_lu(O, threads, res):
    n = LU_N
    t0 = time.time()
    A = posit_matrix(O, LU_SEED, n, n)
    ent = {"n": n, "nb": LU_NB, "seed": LU_SEED, "a": digest(A)}
    ones = np.full(n, ONE, np.uint32)
    b = np.zeros(n, np.uint32)
    assert O.rgemv("N", n, n, ONE, A, ones, 1, 0, b, 1) == 0
    ent["b"] = digest(b)
    rc, ipiv, info = O.rgetrf(A, nb=LU_NB, nthreads=threads)
    assert rc == 0
    ent["factor"] = digests(A)
    ent["ipiv"] = digest(ipiv.astype(np.int32))
    ent["info"] = int(info)
    X = np.asfortranarray(b.reshape(n, 1))
    rc, sinfo = O.rgetrs("N", A, ipiv, X)
    assert rc == 0 and sinfo == 0
    ent["x"] = digest(X[:, 0])
    ent["oracle_s"] = round(time.time() - t0, 1)
    res["lu"] = ent
    print(f"lu: {ent['oracle_s']} s, info {info}", flush=True)
    save(res)


def run_chol(O, threads, res):
    n = CHOL_N
    t0 = time.time()
    S = spd_quantized(O, CHOL_SEED, n)
    ent = {"n": n, "nb": CHOL_NB, "seed": CHOL_SEED, "uplo": "L", "a": digest(S)}
    ones = np.full(n, ONE, np.uint32)
    b = np.zeros(n, np.uint32)
    assert O.rgemv("N", n, n, ONE, S, ones, 1, 0, b, 1) == 0
    ent["b"] = digest(b)
    rc, info = O.rpotrf("L", S, nb=CHOL_NB, nthreads=threads)
    assert rc == 0
    ent["factor"] = digests(S)
    ent["info"] = int(info)
    X = np.asfortranarray(b.reshape(n, 1))
    rc, sinfo = O.rpotrs("L", S, X)
    assert rc == 0 and sinfo == 0
    ent["x"] = digest(X[:, 0])
    ent["oracle_s"] = round(time.time() - t0, 1)
    res["chol"] = ent
    print(f"chol: {ent['oracle_s']} s, info {info}", flush=True)
    save(res)


def save(res):
    tmp = OUT + ".tmp"
    with open(tmp, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, OUT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", choices=["gemm", "lu", "chol"], default=None)
    args = ap.parse_args()
    res = {"gemm": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            res = json.load(f)
        res.setdefault("gemm", {})
    res["generator"] = "tests/golden/gen_fullsize_hashes.py (oracle/liborc.so)"
    O = Oracle()
    if args.only in (None, "chol"):
        run_chol(O, args.threads, res)
    if args.only in (None, "lu"):
        run_lu(O, args.threads, res)
    if args.only in (None, "gemm"):
        run_gemm(O, args.threads, res)


if __name__ == "__main__":
    main()
