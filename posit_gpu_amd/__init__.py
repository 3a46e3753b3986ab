"""posit_gpu_amd -- MI355X (gfx950) posit32 dense linear algebra.

Python access to the C ABI of posit_gpu_amd/lib/libposit_gpu_amd.so (declared in
include/posit_gpu_amd.h).  Every computation runs in the HIP library on the GPU; this
module only moves arrays and calls entry points.  Importing it without the built library
raises ImportError: there is no host fallback.

Posit32 values are numpy uint32 arrays holding the bit patterns; matrices are
column-major (Fortran order), as in the reference's libRpack.

Names that mirror the reference API (RgemmG, Rgemm, ...) take the reference's argument
lists (libRpack/mpblas_posit.h:37-55, libRpack/mplapack_posit.h:38-57) with numpy arrays
in place of posit32 pointers.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libposit_gpu_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"posit_gpu_amd: HIP library not built ({LIB_PATH}); run __graft_entry__.build() "
        "or make -C posit_gpu_amd/csrc"
    )

_lib = ctypes.CDLL(LIB_PATH)

_P = ctypes.c_void_p
_I = ctypes.c_int
_U32 = ctypes.c_uint32
_C = ctypes.c_char
_SZ = ctypes.c_size_t

P32_ZERO = 0x00000000
P32_ONE = 0x40000000
P32_MONE = 0xC0000000
P32_NAR = 0x80000000


class PositError(RuntimeError):
    """A call into the HIP library failed."""

    def __init__(self, status, where, argument=0):
        self.status = status
        self.argument = argument
        msg = _lib.pga_status_string(status).decode()
        if argument:
            msg += f" (parameter {argument})"
        super().__init__(f"{where}: {msg}")


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("compute_units", ctypes.c_int),
        ("clock_khz", ctypes.c_int),
        ("simd_per_cu", ctypes.c_int),
        ("lanes_per_simd", ctypes.c_int),
        ("int32_valu_peak", ctypes.c_double),
        ("hbm_bytes", ctypes.c_size_t),
    ]


def _sig(name, args, res=_I):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("pga_init", [_I])
_sig("pga_finalize", [])
_sig("pga_is_initialized", [])
_sig("pga_status_string", [_I], ctypes.c_char_p)
_sig("pga_illegal_argument", [])
_sig("pga_get_device_info", [ctypes.POINTER(DeviceInfo)])
_sig("pga_set_gemm_impl", [_I])
_sig("pga_malloc", [ctypes.POINTER(_P), _SZ])
_sig("pga_free", [_P])
_sig("pga_memcpy_h2d", [_P, _P, _SZ])
_sig("pga_memcpy_d2h", [_P, _P, _SZ])
_sig("pga_memcpy_d2d", [_P, _P, _SZ])
_sig("pga_memset", [_P, _I, _SZ])
_sig("pga_stream_create", [ctypes.POINTER(_P)])
_sig("pga_stream_destroy", [_P])
_sig("pga_stream_synchronize", [_P])
_sig("pga_device_synchronize", [])
_sig("pga_event_create", [ctypes.POINTER(_P)])
_sig("pga_event_destroy", [_P])
_sig("pga_event_record", [_P, _P])
_sig("pga_event_synchronize", [_P])
_sig("pga_event_elapsed_ms", [ctypes.POINTER(ctypes.c_float), _P, _P])
_sig("pga_rgemm", [_C, _C, _I, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I])
_sig("pga_rgemm_device", [_C, _C, _I, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I, _P])
_sig("pga_rgemm_autotune", [_C, _C, _I, _I, _I, _U32, _U32, ctypes.POINTER(ctypes.c_int)])
_sig("pga_p32_from_f64_device", [_P, _P, _SZ, _P])
_sig("pga_p32_to_f64_device", [_P, _P, _SZ, _P])
_sig("pga_fill_uniform_p32_device", [_P, _SZ, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _P])
_sig("pga_p32_elementwise_device", [_I, _P, _P, _P, _SZ, _P])


def _check(status, where):
    if status != 0:
        raise PositError(status, where, _lib.pga_illegal_argument() if status == 1 else 0)


def init(device=-1):
    """Select the GPU (initOpenCL in the reference, libRpack/Rgemm_opencl.cpp:59)."""
    _check(_lib.pga_init(device), "pga_init")


def initOpenCL():  # noqa: N802 -- reference name (libRpack/def.h:63)
    init(-1)


def finalize():
    _check(_lib.pga_finalize(), "pga_finalize")


def device_info():
    init()
    info = DeviceInfo()
    _check(_lib.pga_get_device_info(ctypes.byref(info)), "pga_get_device_info")
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "compute_units": info.compute_units,
        "clock_khz": info.clock_khz,
        "simd_per_cu": info.simd_per_cu,
        "lanes_per_simd": info.lanes_per_simd,
        "int32_valu_peak": info.int32_valu_peak,
        "hbm_bytes": info.hbm_bytes,
    }


def set_gemm_impl(impl):
    """'tiled' (default) or 'exact' (one element per thread, pattern-level ops)."""
    init()
    _check(_lib.pga_set_gemm_impl({"tiled": 0, "exact": 1}[impl]), "pga_set_gemm_impl")


# ---------------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------------
class DeviceArray:
    """A device allocation holding a numpy-shaped array (column-major for matrices)."""

    def __init__(self, shape, dtype=np.uint32, order="F"):
        init()
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.order = order
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _P()
        _check(_lib.pga_malloc(ctypes.byref(p), max(self.nbytes, 4)), "pga_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.asarray(a)
        order = "F" if a.ndim > 1 else "C"
        d = cls(a.shape, a.dtype, order)
        h = np.asfortranarray(a) if order == "F" else np.ascontiguousarray(a)
        if d.nbytes:
            _check(_lib.pga_memcpy_h2d(d.ptr, h.ctypes.data, d.nbytes), "pga_memcpy_h2d")
        return d

    def to_host(self):
        h = np.empty(self.shape, dtype=self.dtype, order=self.order)
        if self.nbytes:
            _check(_lib.pga_memcpy_d2h(h.ctypes.data, self.ptr, self.nbytes), "pga_memcpy_d2h")
        return h

    def copy_from_host(self, a):
        h = np.asfortranarray(a, dtype=self.dtype) if self.order == "F" else np.ascontiguousarray(a, dtype=self.dtype)
        assert h.nbytes == self.nbytes
        if self.nbytes:
            _check(_lib.pga_memcpy_h2d(self.ptr, h.ctypes.data, self.nbytes), "pga_memcpy_h2d")

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _check(_lib.pga_memcpy_d2d(self.ptr, other.ptr, self.nbytes), "pga_memcpy_d2d")

    def offset(self, elements):
        """Raw device pointer `elements` items into the buffer."""
        return self.ptr + int(elements) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            _lib.pga_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        init()
        p = _P()
        _check(_lib.pga_event_create(ctypes.byref(p)), "pga_event_create")
        self.ptr = p.value

    def record(self, stream=None):
        _check(_lib.pga_event_record(self.ptr, stream), "pga_event_record")

    def synchronize(self):
        _check(_lib.pga_event_synchronize(self.ptr), "pga_event_synchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _check(_lib.pga_event_elapsed_ms(ctypes.byref(ms), self.ptr, later.ptr), "pga_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):
        try:
            _lib.pga_event_destroy(self.ptr)
        except Exception:
            pass


def synchronize(stream=None):
    _check(_lib.pga_stream_synchronize(stream), "pga_stream_synchronize")


# ---------------------------------------------------------------------------------
# codecs and elementwise arithmetic (device)
# ---------------------------------------------------------------------------------
def from_double(x):
    """float64 array -> posit32 patterns, converted on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dx = DeviceArray.from_host(x.ravel())
    dp = DeviceArray(x.size, np.uint32, "C")
    _check(_lib.pga_p32_from_f64_device(dx.ptr, dp.ptr, x.size, None), "pga_p32_from_f64_device")
    synchronize()
    return dp.to_host().reshape(x.shape)


def to_double(p):
    p = np.ascontiguousarray(p, dtype=np.uint32)
    dp = DeviceArray.from_host(p.ravel())
    dx = DeviceArray(p.size, np.float64, "C")
    _check(_lib.pga_p32_to_f64_device(dp.ptr, dx.ptr, p.size, None), "pga_p32_to_f64_device")
    synchronize()
    return dx.to_host().reshape(p.shape)


_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4}


def elementwise(op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint32).ravel()
    b = a if b is None else np.ascontiguousarray(b, dtype=np.uint32).ravel()
    da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
    do = DeviceArray(a.size, np.uint32, "C")
    _check(_lib.pga_p32_elementwise_device(_OPS[op], da.ptr, db.ptr, do.ptr, a.size, None),
           "pga_p32_elementwise_device")
    synchronize()
    return do.to_host()


def fill_uniform(d, seed, lo=-1.0, hi=1.0, stream=None):
    """Fill a DeviceArray of uint32 with posit32(lo + (hi-lo)*u), u from SplitMix64(seed, i)."""
    n = int(np.prod(d.shape))
    _check(_lib.pga_fill_uniform_p32_device(d.ptr, n, seed, lo, hi, stream), "pga_fill_uniform_p32_device")


# ---------------------------------------------------------------------------------
# GEMM
# ---------------------------------------------------------------------------------
def _ch(c):
    return c.encode()[:1] if isinstance(c, str) else bytes(c)[:1]


def rgemm_device(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    """C = alpha*op(A)*op(B) + beta*C on device pointers (ints) or DeviceArrays."""
    pa = a.ptr if isinstance(a, DeviceArray) else a
    pb = b.ptr if isinstance(b, DeviceArray) else b
    pc = c.ptr if isinstance(c, DeviceArray) else c
    _check(_lib.pga_rgemm_device(_ch(transa), _ch(transb), m, n, k, alpha, pa, lda, pb, ldb, beta, pc, ldc,
                                 stream), "pga_rgemm_device")


def rgemm_autotune(transa, transb, m, n, k, alpha=0xC0000000, beta=0x40000000):
    """Times the tile edges 64 / 32 / 16 for this exact shape on scratch operands and
    records the fastest for later GEMMs of that shape; returns it."""
    t = ctypes.c_int()
    _check(_lib.pga_rgemm_autotune(_ch(transa), _ch(transb), m, n, k, alpha, beta, ctypes.byref(t)),
           "pga_rgemm_autotune")
    return t.value


def RgemmG(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):  # noqa: N802
    """libRpack/mpblas_posit.h:49-55 on host numpy uint32 buffers (C updated in place).

    A, B and C are flat (or Fortran-ordered) buffers addressed with the leading
    dimensions, exactly as the reference's posit32 pointers.
    """
    for x in (A, B, C):
        assert x.dtype == np.uint32
    assert C.flags.writeable and (C.flags.f_contiguous or C.flags.c_contiguous)
    _check(_lib.pga_rgemm(_ch(transa), _ch(transb), m, n, k, alpha, A.ctypes.data, lda, B.ctypes.data, ldb,
                          beta, C.ctypes.data, ldc), "pga_rgemm")


Rgemm = RgemmG  # both run on the GPU with Rgemm's per-element order


def rgemm(transa, transb, alpha, A, B, beta, C):
    """Convenience form on Fortran-ordered uint32 matrices; returns the new C."""
    A = np.asfortranarray(A, dtype=np.uint32)
    B = np.asfortranarray(B, dtype=np.uint32)
    C = np.array(C, dtype=np.uint32, order="F", copy=True)
    m, n = C.shape
    k = A.shape[0] if transa.upper() in "TC" else A.shape[1]
    RgemmG(transa, transb, m, n, k, alpha, A, max(1, A.shape[0]), B, max(1, B.shape[0]), beta, C, max(1, m))
    return C


# ---------------------------------------------------------------------------------
# BLAS-2/3 and LAPACK (host-pointer drop-ins and device forms)
# ---------------------------------------------------------------------------------
_IP = ctypes.POINTER(ctypes.c_int)
_sig("pga_rtrsm", [_C, _C, _C, _C, _I, _I, _U32, _P, _I, _P, _I])
_sig("pga_rtrsm_device", [_C, _C, _C, _C, _I, _I, _U32, _P, _I, _P, _I, _P])
_sig("pga_rsyrk", [_C, _C, _I, _I, _U32, _P, _I, _U32, _P, _I])
_sig("pga_rsyrk_device", [_C, _C, _I, _I, _U32, _P, _I, _U32, _P, _I, _P])
_sig("pga_rgemv", [_C, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I])
_sig("pga_rgemv_device", [_C, _I, _I, _U32, _P, _I, _P, _I, _U32, _P, _I, _P])
_sig("pga_rscal", [_I, _U32, _P, _I])
_sig("pga_rscal_device", [_I, _U32, _P, _I, _P])
_sig("pga_iramax", [_I, _P, _I, _IP])
_sig("pga_iramax_device", [_I, _P, _I, _P, _P])
_sig("pga_rlaswp", [_I, _P, _I, _I, _I, _P, _I])
_sig("pga_rlaswp_device", [_I, _P, _I, _I, _I, _P, _I, _P])
_sig("pga_rgetrf", [_I, _I, _P, _I, _P, _IP, _I])
_sig("pga_rgetrf_device", [_I, _I, _P, _I, _P, _P, _I, _P])
_sig("pga_rgetrf2", [_I, _I, _P, _I, _P, _IP])
_sig("pga_rgetrf2_device", [_I, _I, _P, _I, _P, _P, _P])
_sig("pga_rgetrs", [_C, _I, _I, _P, _I, _P, _P, _I, _IP])
_sig("pga_rgetrs_device", [_C, _I, _I, _P, _I, _P, _P, _I, _P])
_sig("pga_rpotrf", [_C, _I, _P, _I, _IP, _I])
_sig("pga_rpotrf_device", [_C, _I, _P, _I, _P, _I, _P])
_sig("pga_rpotrf2", [_C, _I, _P, _I, _IP])
_sig("pga_rpotrf2_device", [_C, _I, _P, _I, _P, _P])
_sig("pga_rpotrs", [_C, _I, _I, _P, _I, _P, _I, _IP])
_sig("pga_rpotrs_device", [_C, _I, _I, _P, _I, _P, _I, _P])


def _buf(x, dtype=np.uint32):
    assert x.dtype == dtype, x.dtype
    assert x.flags.f_contiguous or x.flags.c_contiguous
    return x.ctypes.data


def _ld(x):
    return max(1, x.shape[0]) if x.ndim > 1 else max(1, x.size)


def Rtrsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):  # noqa: N802
    """libRpack/mpblas_posit.h:38 (Rtrsm.cpp:31-291), on the GPU; b updated in place."""
    _check(_lib.pga_rtrsm(_ch(side), _ch(uplo), _ch(transa), _ch(diag), m, n, alpha, _buf(a), lda, _buf(b), ldb),
           "pga_rtrsm")


def Rsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc):  # noqa: N802
    """libRpack/mpblas_posit.h:39 (Rsyrk.cpp:31-215)."""
    _check(_lib.pga_rsyrk(_ch(uplo), _ch(trans), n, k, alpha, _buf(a), lda, beta, _buf(c), ldc), "pga_rsyrk")


def Rgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):  # noqa: N802
    """libRpack/mpblas_posit.h:40 (Rgemv.cpp:31-204)."""
    _check(_lib.pga_rgemv(_ch(trans), m, n, alpha, _buf(a), lda, _buf(x), incx, beta, _buf(y), incy), "pga_rgemv")


def Rscal(n, da, dx, incx):  # noqa: N802
    """libRpack/mpblas_posit.h:43 (Rscal.cpp:62-119)."""
    _check(_lib.pga_rscal(n, da, _buf(dx), incx), "pga_rscal")


def iRamax(n, dx, incx):  # noqa: N802
    """libRpack/mpblas_posit.h:44 (iRamax.cpp:31-88): 1-based index."""
    r = ctypes.c_int()
    _check(_lib.pga_iramax(n, _buf(dx), incx, ctypes.byref(r)), "pga_iramax")
    return r.value


def Rlaswp(n, a, lda, k1, k2, ipiv, incx):  # noqa: N802
    """libRpack/mplapack_posit.h:54 (Rlaswp.cpp:32-111); ipiv int32, 1-based."""
    _check(_lib.pga_rlaswp(n, _buf(a), lda, k1, k2, _buf(ipiv, np.int32), incx), "pga_rlaswp")


def Rgetrf_opencl(m, n, a, lda, ipiv, nbb=0):  # noqa: N802
    """libRpack/mplapack_posit.h:49 (Rgetrf_opencl.cpp:32-154); returns info."""
    info = ctypes.c_int()
    _check(_lib.pga_rgetrf(m, n, _buf(a), lda, _buf(ipiv, np.int32), ctypes.byref(info), nbb), "pga_rgetrf")
    return info.value


def Rgetrf(m, n, a, lda, ipiv):  # noqa: N802
    """libRpack/mplapack_posit.h:48 (Rgetrf.cpp): nb from iMlaenv (64)."""
    return Rgetrf_opencl(m, n, a, lda, ipiv, 0)


def Rgetrf2(m, n, a, lda, ipiv):  # noqa: N802
    """libRpack/mplapack_posit.h:51 (Rgetrf2.cpp:32-185)."""
    info = ctypes.c_int()
    _check(_lib.pga_rgetrf2(m, n, _buf(a), lda, _buf(ipiv, np.int32), ctypes.byref(info)), "pga_rgetrf2")
    return info.value


Rgetrf2_opencl = Rgetrf2


def Rgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb):  # noqa: N802
    """libRpack/mplapack_posit.h:46 (Rgetrs.cpp:32-118); returns info."""
    info = ctypes.c_int()
    _check(_lib.pga_rgetrs(_ch(trans), n, nrhs, _buf(a), lda, _buf(ipiv, np.int32), _buf(b), ldb, ctypes.byref(info)),
           "pga_rgetrs")
    return info.value


def Rpotrf_opencl(uplo, n, a, lda, nb=0):  # noqa: N802
    """libRpack/mplapack_posit.h:43 (Rpotrf_opencl.cpp:32-166); returns info."""
    info = ctypes.c_int()
    _check(_lib.pga_rpotrf(_ch(uplo), n, _buf(a), lda, ctypes.byref(info), nb), "pga_rpotrf")
    return info.value


def Rpotrf(uplo, n, a, lda):  # noqa: N802
    return Rpotrf_opencl(uplo, n, a, lda, 0)


def Rpotrf2(uplo, n, a, lda):  # noqa: N802
    info = ctypes.c_int()
    _check(_lib.pga_rpotrf2(_ch(uplo), n, _buf(a), lda, ctypes.byref(info)), "pga_rpotrf2")
    return info.value


def Rpotrs(uplo, n, nrhs, a, lda, b, ldb):  # noqa: N802
    info = ctypes.c_int()
    _check(_lib.pga_rpotrs(_ch(uplo), n, nrhs, _buf(a), lda, _buf(b), ldb, ctypes.byref(info)), "pga_rpotrs")
    return info.value


def _dp(x):
    return x.ptr if isinstance(x, DeviceArray) else x


def getrf_device(m, n, a, lda, ipiv, info, nb=0, stream=None):
    _check(_lib.pga_rgetrf_device(m, n, _dp(a), lda, _dp(ipiv), _dp(info), nb, stream), "pga_rgetrf_device")


def getrs_device(trans, n, nrhs, a, lda, ipiv, b, ldb, stream=None):
    _check(_lib.pga_rgetrs_device(_ch(trans), n, nrhs, _dp(a), lda, _dp(ipiv), _dp(b), ldb, stream),
           "pga_rgetrs_device")


def potrf_device(uplo, n, a, lda, info, nb=0, stream=None):
    _check(_lib.pga_rpotrf_device(_ch(uplo), n, _dp(a), lda, _dp(info), nb, stream), "pga_rpotrf_device")


def potrs_device(uplo, n, nrhs, a, lda, b, ldb, stream=None):
    _check(_lib.pga_rpotrs_device(_ch(uplo), n, nrhs, _dp(a), lda, _dp(b), ldb, stream), "pga_rpotrs_device")


def gemv_device(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream=None):
    _check(_lib.pga_rgemv_device(_ch(trans), m, n, alpha, _dp(a), lda, _dp(x), incx, beta, _dp(y), incy, stream),
           "pga_rgemv_device")


_sig("pga_backward_error_device", [_I, _P, _I, _P, _P, ctypes.POINTER(ctypes.c_double), _P])


def backward_error_device(n, a, lda, x, b, stream=None):
    """||A x - b||_inf / (||A||_inf ||x||_inf) in FP64 on the GPU (posit32 device arrays)."""
    r = ctypes.c_double()
    _check(_lib.pga_backward_error_device(n, _dp(a), lda, _dp(x), _dp(b), ctypes.byref(r), stream),
           "pga_backward_error_device")
    return r.value


# ---------------------------------------------------------------------------------
# batched GEMM plans
# ---------------------------------------------------------------------------------
_PP = ctypes.POINTER(_P)
_sig("pga_batch_plan_create", [ctypes.POINTER(_P), _C, _C, _I, _IP, _IP, _IP, ctypes.POINTER(_U32), _PP, _IP,
                               _PP, _IP, ctypes.POINTER(_U32), _PP, _IP, _I, _I])
_sig("pga_batch_plan_execute", [_P, _P])
_sig("pga_batch_plan_capture_graph", [_P, _P])
_sig("pga_batch_plan_launch_graph", [_P, _P])
_sig("pga_batch_plan_info", [_P, _IP, _IP, _IP, ctypes.POINTER(ctypes.c_double)])
_sig("pga_batch_plan_schedule", [_P, _IP])
_sig("pga_batch_plan_destroy", [_P])


class BatchPlan:
    """Batched GEMM C_i = alpha_i*op(A_i)*op(B_i) + beta_i*C_i over device buffers.

    problems: list of dicts with keys m, n, k, alpha, beta, a, lda, b, ldb, c, ldc where
    a, b, c are device pointers (ints) or DeviceArrays.  tile=0 autotunes."""

    def __init__(self, transa, transb, problems, tile=0, grid=0):
        init()
        cnt = len(problems)
        ia = lambda key: (ctypes.c_int * max(1, cnt))(*[int(p[key]) for p in problems])  # noqa: E731
        ua = lambda key: (_U32 * max(1, cnt))(*[int(p[key]) for p in problems])  # noqa: E731
        pa = lambda key: (_P * max(1, cnt))(*[_dp(p[key]) for p in problems])  # noqa: E731
        self._keep = [ia("m"), ia("n"), ia("k"), ua("alpha"), pa("a"), ia("lda"), pa("b"), ia("ldb"), ua("beta"),
                      pa("c"), ia("ldc")]
        m, n, k, al, a, lda, b, ldb, be, c, ldc = self._keep
        h = _P()
        _check(_lib.pga_batch_plan_create(ctypes.byref(h), _ch(transa), _ch(transb), cnt, m, n, k, al, a, lda, b,
                                          ldb, be, c, ldc, tile, grid), "pga_batch_plan_create")
        self.h = h.value

    def execute(self, stream=None):
        _check(_lib.pga_batch_plan_execute(self.h, stream), "pga_batch_plan_execute")

    def capture_graph(self, stream=None):
        _check(_lib.pga_batch_plan_capture_graph(self.h, stream), "pga_batch_plan_capture_graph")

    def launch_graph(self, stream=None):
        _check(_lib.pga_batch_plan_launch_graph(self.h, stream), "pga_batch_plan_launch_graph")

    def info(self):
        t, g, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        _check(_lib.pga_batch_plan_info(self.h, ctypes.byref(t), ctypes.byref(g), ctypes.byref(n), ctypes.byref(ms)),
               "pga_batch_plan_info")
        sc = ctypes.c_int()
        _check(_lib.pga_batch_plan_schedule(self.h, ctypes.byref(sc)), "pga_batch_plan_schedule")
        return {"tile": t.value, "grid": g.value, "total_tiles": n.value, "tune_ms": ms.value, "schedule": sc.value}

    def __del__(self):
        try:
            _lib.pga_batch_plan_destroy(self.h)
        except Exception:
            pass


NOALLOC_MAX_PROBLEMS = 4096
_sig("pga_batch_gemm_noalloc", [_C, _C, _I, _IP, _IP, _IP, ctypes.POINTER(_U32), _PP, _IP, _PP, _IP,
                                ctypes.POINTER(_U32), _PP, _IP, _I, _I, _P])


def batch_gemm_noalloc(transa, transb, problems, tile=0, grid=0, stream=None):
    """pga_batch_gemm_noalloc: the batched GEMM of BatchPlan in one call that allocates no
    device memory (descriptors in the library's fixed device array, at most
    NOALLOC_MAX_PROBLEMS problems).  Asynchronous on `stream`; calls that overlap on one
    GPU race."""
    init()
    cnt = len(problems)
    ia = lambda key: (ctypes.c_int * max(1, cnt))(*[int(p[key]) for p in problems])  # noqa: E731
    ua = lambda key: (_U32 * max(1, cnt))(*[int(p[key]) for p in problems])  # noqa: E731
    pa = lambda key: (_P * max(1, cnt))(*[_dp(p[key]) for p in problems])  # noqa: E731
    _check(_lib.pga_batch_gemm_noalloc(_ch(transa), _ch(transb), cnt, ia("m"), ia("n"), ia("k"), ua("alpha"),
                                       pa("a"), ia("lda"), pa("b"), ia("ldb"), ua("beta"), pa("c"), ia("ldc"),
                                       tile, grid, stream), "pga_batch_gemm_noalloc")


# ---------------------------------------------------------------------------------
# batched factorizations / solves
# ---------------------------------------------------------------------------------
TASK_GETRF, TASK_POTRF, TASK_GETRS, TASK_POTRS, TASK_GESV, TASK_POSV = range(6)
FACTOR_TASK_MAX_N = 512        # one workgroup per task up to this order
BLOCKED_TASK_MAX_N = 16384     # larger tasks advance together one column block at a time
_TASK_KINDS = {"getrf": TASK_GETRF, "potrf": TASK_POTRF, "getrs": TASK_GETRS, "potrs": TASK_POTRS,
               "gesv": TASK_GESV, "posv": TASK_POSV}


class FactorTask(ctypes.Structure):
    """pga_factor_task of include/posit_gpu_amd.h."""
    _fields_ = [("kind", ctypes.c_int), ("uplo", ctypes.c_char), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("nb", ctypes.c_int), ("a", _P), ("lda", ctypes.c_int), ("ipiv", _P), ("b", _P), ("ldb", ctypes.c_int),
                ("info", _P)]


_sig("pga_factor_plan_create", [ctypes.POINTER(_P), _I, ctypes.POINTER(FactorTask), _I])
_sig("pga_factor_plan_execute", [_P, _P])
_sig("pga_factor_plan_capture_graph", [_P, _P])
_sig("pga_factor_plan_launch_graph", [_P, _P])
_sig("pga_factor_plan_info", [_P, _IP, _IP])
_sig("pga_factor_plan_autotune", [_P, _IP, _IP, ctypes.POINTER(ctypes.c_double)])
_sig("pga_factor_plan_destroy", [_P])


class FactorPlan:
    """Many independent factorizations / solves: orders up to FACTOR_TASK_MAX_N in one
    persistent launch (a workgroup per task), larger ones (up to BLOCKED_TASK_MAX_N)
    together one column block at a time.

    tasks: list of dicts with keys kind ("getrf", "potrf", "getrs", "potrs", "gesv",
    "posv" or the TASK_* value), n, a, lda and, as the kind needs, uplo, nb, ipiv, b,
    ldb, nrhs, info; a, ipiv, b, info are DeviceArrays or device pointers."""

    def __init__(self, tasks, grid=0):
        init()
        cnt = len(tasks)
        arr = (FactorTask * max(1, cnt))()
        for i, t in enumerate(tasks):
            kind = t["kind"]
            arr[i].kind = _TASK_KINDS[kind] if isinstance(kind, str) else int(kind)
            arr[i].uplo = _ch(t.get("uplo", "L"))
            arr[i].n = int(t["n"])
            arr[i].nrhs = int(t.get("nrhs", 0))
            arr[i].nb = int(t.get("nb", 0))
            arr[i].a = _dp(t["a"])
            arr[i].lda = int(t.get("lda", max(1, t["n"])))
            arr[i].ipiv = _dp(t.get("ipiv")) if t.get("ipiv") is not None else None
            arr[i].b = _dp(t.get("b")) if t.get("b") is not None else None
            arr[i].ldb = int(t.get("ldb", max(1, t["n"])))
            arr[i].info = _dp(t.get("info")) if t.get("info") is not None else None
        h = _P()
        _check(_lib.pga_factor_plan_create(ctypes.byref(h), cnt, arr, grid), "pga_factor_plan_create")
        self.h = h.value

    def execute(self, stream=None):
        _check(_lib.pga_factor_plan_execute(self.h, stream), "pga_factor_plan_execute")

    def capture_graph(self, stream=None):
        _check(_lib.pga_factor_plan_capture_graph(self.h, stream), "pga_factor_plan_capture_graph")

    def launch_graph(self, stream=None):
        _check(_lib.pga_factor_plan_launch_graph(self.h, stream), "pga_factor_plan_launch_graph")

    def info(self):
        c, g = ctypes.c_int(), ctypes.c_int()
        _check(_lib.pga_factor_plan_info(self.h, ctypes.byref(c), ctypes.byref(g)), "pga_factor_plan_info")
        return {"count": c.value, "grid": g.value}

    def autotune(self):
        """Tunes the persistent launch (workgroups per CU, task order) on scratch copies."""
        g, sc, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(_lib.pga_factor_plan_autotune(self.h, ctypes.byref(g), ctypes.byref(sc), ctypes.byref(ms)),
               "pga_factor_plan_autotune")
        return {"grid": g.value, "schedule": sc.value, "tune_ms": ms.value}

    def __del__(self):
        try:
            _lib.pga_factor_plan_destroy(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------------------------
# serving (continuous batching of solve requests)
# ---------------------------------------------------------------------------------
SERVER_BATCHED, SERVER_PLAIN = 0, 1
REQ_LU_SOLVE, REQ_CHOL_SOLVE, REQ_RESOLVE = 0, 1, 2
REQ_OK, REQ_NOT_RESIDENT, REQ_TOO_LARGE, REQ_FAILED = 0, 1, 2, 3


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _SZ), ("staging_bytes", _SZ), ("mode", ctypes.c_int),
                ("batch_max_n", ctypes.c_int), ("max_batch", ctypes.c_int), ("batch_slots", ctypes.c_int),
                ("workers", ctypes.c_int), ("chol_uplo", ctypes.c_char)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("nrhs", ctypes.c_int), ("a", _P), ("b", _P),
                ("x", _P), ("factor", ctypes.c_int64), ("keep_factor", ctypes.c_int)]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("factor", ctypes.c_int64), ("status", ctypes.c_int), ("info", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("solved", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("batched_requests", ctypes.c_int64), ("evictions", ctypes.c_int64),
                ("resident_factors", ctypes.c_int), ("bytes_in_use", _SZ), ("bytes_peak", _SZ)]


_sig("pga_server_create", [ctypes.POINTER(_P), ctypes.POINTER(ServerConfig)])
_sig("pga_server_submit", [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)])
_sig("pga_server_wait", [_P, ctypes.c_int64, ctypes.POINTER(RequestResult)])
_sig("pga_server_drain", [_P])
_sig("pga_server_release_factor", [_P, ctypes.c_int64])
_sig("pga_server_get_stats", [_P, ctypes.POINTER(ServerStats)])
_sig("pga_server_destroy", [_P])


class Server:
    """Posit32 solve requests served with continuous batching (see pga_server_* in
    include/posit_gpu_amd.h).  Matrices are host uint32 arrays of posit patterns."""

    def __init__(self, hbm_budget_bytes, mode="batched", batch_max_n=512, max_batch=1024, batch_slots=2, workers=3,
                 chol_uplo="U", staging_bytes=0):
        init()
        cfg = ServerConfig(hbm_budget_bytes, staging_bytes, SERVER_PLAIN if mode == "plain" else SERVER_BATCHED,
                           batch_max_n, max_batch, batch_slots, workers, _ch(chol_uplo))
        h = _P()
        _check(_lib.pga_server_create(ctypes.byref(h), ctypes.byref(cfg)), "pga_server_create")
        self.h = h.value
        self._live = {}

    def submit(self, kind, n, a=None, b=None, factor=0, keep_factor=True):
        """kind: "lu", "chol" (factor and solve) or "resolve" (against factor id)."""
        kinds = {"lu": REQ_LU_SOLVE, "chol": REQ_CHOL_SOLVE, "resolve": REQ_RESOLVE}
        a = None if a is None else np.asfortranarray(a, dtype=np.uint32)
        b = None if b is None else np.asfortranarray(b, dtype=np.uint32).reshape(n, -1, order="F")
        nrhs = 0 if b is None else b.shape[1]
        x = np.zeros((n, max(nrhs, 1)), np.uint32, order="F")
        r = Request(kinds[kind], n, nrhs, None if a is None else a.ctypes.data, None if b is None else b.ctypes.data,
                    x.ctypes.data, int(factor), 1 if keep_factor else 0)
        rid = ctypes.c_int64()
        _check(_lib.pga_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)), "pga_server_submit")
        self._live[rid.value] = (a, b, x, nrhs)
        return rid.value

    def wait(self, rid):
        res = RequestResult()
        _check(_lib.pga_server_wait(self.h, rid, ctypes.byref(res)), "pga_server_wait")
        a, b, x, nrhs = self._live.pop(rid)
        return {"id": res.id, "factor": res.factor, "status": res.status, "info": res.info,
                "t_submit": res.t_submit, "t_admit": res.t_admit, "t_done": res.t_done,
                "x": x[:, :nrhs] if nrhs else None}

    def drain(self):
        _check(_lib.pga_server_drain(self.h), "pga_server_drain")

    def release_factor(self, factor):
        _check(_lib.pga_server_release_factor(self.h, int(factor)), "pga_server_release_factor")

    def stats(self):
        st = ServerStats()
        _check(_lib.pga_server_get_stats(self.h, ctypes.byref(st)), "pga_server_get_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            _lib.pga_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
