"""Results do not depend on what else runs on the device: LU and Cholesky factorizations on
one stream while another stream runs large GEMMs equal the results computed on an
otherwise idle device, bit for bit; so do LU factorizations run concurrently from several
host threads on their own streams."""
import ctypes
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ONE = 0x40000000


def _stream(pga):
    p = ctypes.c_void_p()
    pga._check(pga._lib.pga_stream_create(ctypes.byref(p)), "pga_stream_create")
    return p.value


def _rand(pga, m, n, seed):
    d = pga.DeviceArray((m, n))
    pga.fill_uniform(d, seed)
    pga.synchronize()
    return d.to_host()


def _lu(pga, M, nb, s):
    n = M.shape[0]
    A = pga.DeviceArray.from_host(M)
    ip = pga.DeviceArray(n, np.int32, "C")
    inf = pga.DeviceArray(1, np.int32, "C")
    pga.getrf_device(n, n, A, n, ip, inf, nb, s)
    pga.synchronize(s)
    return A.to_host(), ip.to_host(), int(inf.to_host()[0])


def _panel(pga, P, s):
    m, w = P.shape
    A = pga.DeviceArray.from_host(P)
    ip = pga.DeviceArray(w, np.int32, "C")
    inf = pga.DeviceArray(1, np.int32, "C")
    pga._check(pga._lib.pga_rgetrf2_device(m, w, A.ptr, m, ip.ptr, inf.ptr, s), "pga_rgetrf2_device")
    pga.synchronize(s)
    return A.to_host(), ip.to_host(), int(inf.to_host()[0])


def _chol(pga, M, uplo, s):
    n = M.shape[0]
    Md = pga.DeviceArray.from_host(M)
    A = pga.DeviceArray((n, n))
    inf = pga.DeviceArray(1, np.int32, "C")
    pga.rgemm_device("N", "T", n, n, n, ONE, Md, n, Md, n, 0, A, n, s)
    pga.potrf_device(uplo, n, A, n, inf, 64, s)
    pga.synchronize(s)
    return A.to_host(), int(inf.to_host()[0])


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_factorizations_under_concurrent_gemm(pga):
    M = _rand(pga, 400, 400, 70)
    P = _rand(pga, 600, 64, 71)
    cases = [("getrf nb=64", lambda s: _lu(pga, M, 64, s)), ("getrf2 panel 600x64", lambda s: _panel(pga, P, s)),
             ("potrf U", lambda s: _chol(pga, M, "U", s)), ("potrf L", lambda s: _chol(pga, M, "L", s))]
    idle = {name: f(None) for name, f in cases}
    N = 2048
    GA = pga.DeviceArray((N, N))
    pga.fill_uniform(GA, 11)
    GC = pga.DeviceArray((N, N))
    sb, sl = _stream(pga), _stream(pga)
    stop = threading.Event()

    def background():
        while not stop.is_set():
            pga.rgemm_device("N", "N", N, N, N, ONE, GA, N, GA, N, 0, GC, N, sb)
            pga.synchronize(sb)

    th = threading.Thread(target=background)
    th.start()
    time.sleep(0.2)
    bad = []
    try:
        for rep in range(2):
            for name, f in cases:
                if not _same(f(sl), idle[name]):
                    bad.append((rep, name))
    finally:
        stop.set()
        th.join()
    assert not bad, bad


def test_lu_from_concurrent_threads(pga):
    mats = [_rand(pga, n, n, 80 + i) for i, n in enumerate([317, 334, 392, 413])]
    seq = [_lu(pga, M, 64, None) for M in mats]
    streams = [_stream(pga) for _ in mats]
    for rep in range(2):
        out = [None] * len(mats)

        def work(i):
            out[i] = _lu(pga, mats[i], 64, streams[i])

        ts = [threading.Thread(target=work, args=(i,)) for i in range(len(mats))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert all(_same(o, q) for o, q in zip(out, seq)), rep
