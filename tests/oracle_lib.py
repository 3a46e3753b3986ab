"""ctypes access to the CPU oracle (oracle/liborc.so) and to the host build of the
reference's own posit32 add/mul (oracle/_ref/libposit_ref.so).

Test infrastructure only: the product package never imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PGA_ORACLE_LIB names another build of the oracle (the mutation tests load planted-mistake
# copies this way)
ORC_PATH = os.environ.get("PGA_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libposit_ref.so")

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_C = ctypes.c_char


def _ptr(a):
    return _P(a.ctypes.data)


class Oracle:
    def __init__(self, path=ORC_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        self.L = L
        for name in ("orc_p32_add", "orc_p32_sub", "orc_p32_mul", "orc_p32_div"):
            f = getattr(L, name)
            f.restype = _U32
            f.argtypes = [_U32, _U32]
        for name in ("orc_p32_sqrt", "orc_p32_neg", "orc_p32_abs"):
            f = getattr(L, name)
            f.restype = _U32
            f.argtypes = [_U32]
        L.orc_p32_from_double.restype = _U32
        L.orc_p32_from_double.argtypes = [ctypes.c_double]
        L.orc_p32_to_double.restype = ctypes.c_double
        L.orc_p32_to_double.argtypes = [_U32]
        L.orc_rgemm.restype = _I
        L.orc_rgemm.argtypes = [_C, _C, _I, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I, _I]
        L.orc_rtrsm.restype = _I
        L.orc_rtrsm.argtypes = [_C, _C, _C, _C, _I, _I, _U32, _P, _I, _P, _I]
        L.orc_rsyrk.restype = _I
        L.orc_rsyrk.argtypes = [_C, _C, _I, _I, _U32, _P, _I, _U32, _P, _I]
        L.orc_rgemv.restype = _I
        L.orc_rgemv.argtypes = [_C, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I]
        L.orc_rscal.restype = None
        L.orc_rscal.argtypes = [_I, _U32, _P, _I]
        L.orc_iramax.restype = _I
        L.orc_iramax.argtypes = [_I, _P, _I]
        L.orc_rlaswp.restype = None
        L.orc_rlaswp.argtypes = [_I, _P, _I, _I, _I, _P, _I]
        L.orc_rgetrf2.restype = _I
        L.orc_rgetrf2.argtypes = [_I, _I, _P, _I, _P, _P, _I]
        L.orc_rgetrf.restype = _I
        L.orc_rgetrf.argtypes = [_I, _I, _P, _I, _P, _P, _I, _I]
        L.orc_rgetrs.restype = _I
        L.orc_rgetrs.argtypes = [_C, _I, _I, _P, _I, _P, _P, _I, _P]
        L.orc_rpotrf2.restype = _I
        L.orc_rpotrf2.argtypes = [_C, _I, _P, _I, _P]
        L.orc_rpotrf.restype = _I
        L.orc_rpotrf.argtypes = [_C, _I, _P, _I, _P, _I, _I]
        L.orc_rpotrs.restype = _I
        L.orc_rpotrs.argtypes = [_C, _I, _I, _P, _I, _P, _I, _P]

    # ---- scalar / elementwise ----
    def add(self, a, b):
        return int(self.L.orc_p32_add(a, b))

    def sub(self, a, b):
        return int(self.L.orc_p32_sub(a, b))

    def mul(self, a, b):
        return int(self.L.orc_p32_mul(a, b))

    def div(self, a, b):
        return int(self.L.orc_p32_div(a, b))

    def sqrt(self, a):
        return int(self.L.orc_p32_sqrt(a))

    def from_double(self, x):
        return int(self.L.orc_p32_from_double(float(x)))

    def to_double(self, a):
        return float(self.L.orc_p32_to_double(a))

    def _binv(self, name, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.empty_like(a)
        getattr(self.L, name)(_ptr(a), _ptr(b), _ptr(out), ctypes.c_size_t(a.size))
        return out

    def add_v(self, a, b):
        return self._binv("orc_p32_add_v", a, b)

    def sub_v(self, a, b):
        return self._binv("orc_p32_sub_v", a, b)

    def mul_v(self, a, b):
        return self._binv("orc_p32_mul_v", a, b)

    def div_v(self, a, b):
        return self._binv("orc_p32_div_v", a, b)

    def sqrt_v(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.empty_like(a)
        self.L.orc_p32_sqrt_v(_ptr(a), _ptr(out), ctypes.c_size_t(a.size))
        return out

    def from_double_v(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(x.shape, dtype=np.uint32)
        self.L.orc_p32_from_double_v(_ptr(x), _ptr(out), ctypes.c_size_t(x.size))
        return out

    def to_double_v(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.empty(a.shape, dtype=np.float64)
        self.L.orc_p32_to_double_v(_ptr(a), _ptr(out), ctypes.c_size_t(a.size))
        return out

    # ---- BLAS / LAPACK on column-major numpy arrays (Fortran order, uint32) ----
    @staticmethod
    def _ld(x):
        assert x.dtype == np.uint32 and x.flags.f_contiguous
        return max(1, x.shape[0])

    def rgemm(self, ta, tb, m, n, k, alpha, a, b, beta, c, nthreads=8, lda=None, ldb=None, ldc=None):
        lda = lda or self._ld(a)
        ldb = ldb or self._ld(b)
        ldc = ldc or self._ld(c)
        return self.L.orc_rgemm(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb,
                                beta, _ptr(c), ldc, nthreads)

    def rtrsm(self, side, uplo, transa, diag, m, n, alpha, a, b):
        return self.L.orc_rtrsm(side.encode(), uplo.encode(), transa.encode(), diag.encode(), m, n, alpha,
                                _ptr(a), self._ld(a), _ptr(b), self._ld(b))

    def rsyrk(self, uplo, trans, n, k, alpha, a, beta, c):
        return self.L.orc_rsyrk(uplo.encode(), trans.encode(), n, k, alpha, _ptr(a), self._ld(a), beta,
                                _ptr(c), self._ld(c))

    def rgemv(self, trans, m, n, alpha, a, x, incx, beta, y, incy):
        return self.L.orc_rgemv(trans.encode(), m, n, alpha, _ptr(a), self._ld(a), _ptr(x), incx, beta,
                                _ptr(y), incy)

    def iramax(self, x, incx=1, n=None):
        n = x.size // max(1, incx) if n is None else n
        return self.L.orc_iramax(n, _ptr(x), incx)

    def rgetrf(self, a, nb=0, nthreads=8):
        m, n = a.shape
        ipiv = np.zeros(min(m, n), dtype=np.int32)
        info = np.zeros(1, dtype=np.int32)
        rc = self.L.orc_rgetrf(m, n, _ptr(a), self._ld(a), _ptr(ipiv), _ptr(info), nb, nthreads)
        return rc, ipiv, int(info[0])

    def rgetrf2(self, a, nthreads=8):
        m, n = a.shape
        ipiv = np.zeros(min(m, n), dtype=np.int32)
        info = np.zeros(1, dtype=np.int32)
        rc = self.L.orc_rgetrf2(m, n, _ptr(a), self._ld(a), _ptr(ipiv), _ptr(info), nthreads)
        return rc, ipiv, int(info[0])

    def rgetrs(self, trans, a, ipiv, b):
        n = a.shape[0]
        info = np.zeros(1, dtype=np.int32)
        ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
        rc = self.L.orc_rgetrs(trans.encode(), n, b.shape[1], _ptr(a), self._ld(a), _ptr(ipiv), _ptr(b),
                               self._ld(b), _ptr(info))
        return rc, int(info[0])

    def rpotrf(self, uplo, a, nb=0, nthreads=8):
        n = a.shape[0]
        info = np.zeros(1, dtype=np.int32)
        rc = self.L.orc_rpotrf(uplo.encode(), n, _ptr(a), self._ld(a), _ptr(info), nb, nthreads)
        return rc, int(info[0])

    def rpotrf2(self, uplo, a):
        n = a.shape[0]
        info = np.zeros(1, dtype=np.int32)
        rc = self.L.orc_rpotrf2(uplo.encode(), n, _ptr(a), self._ld(a), _ptr(info))
        return rc, int(info[0])

    def rpotrs(self, uplo, a, b):
        n = a.shape[0]
        info = np.zeros(1, dtype=np.int32)
        rc = self.L.orc_rpotrs(uplo.encode(), n, b.shape[1], _ptr(a), self._ld(a), _ptr(b), self._ld(b),
                               _ptr(info))
        return rc, int(info[0])


class RefPosit:
    """The reference's own p32_add / p32_mul (kernel_entry.cl), built for the host."""

    def __init__(self, path=REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)

    def _binv(self, name, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.empty_like(a)
        getattr(self.L, name)(_ptr(a), _ptr(b), _ptr(out), ctypes.c_size_t(a.size))
        return out

    def add_v(self, a, b):
        return self._binv("ref_p32_add_v", a, b)

    def mul_v(self, a, b):
        return self._binv("ref_p32_mul_v", a, b)


def ref_available():
    return os.path.exists(REF_PATH)
