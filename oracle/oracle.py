"""TEST INFRASTRUCTURE ONLY: Python access to the oracle.

The oracle is a CPU restatement of HPAC/TTC's generated scalar reference
loop nest (ttc/transpose.py:126-156, 394-473, 570-573 as configured by
ttc/transposeGenerator.py:174-180); see ttc_oracle.cpp for the per-function
citations.  Only tests/, the package smoke check and bench.py's CPU-baseline
leg may import this module.  The product (ttc_amd) never does.

Arrays are numpy arrays holding the *flat memory* of column-major tensors
(index 0 is stride-1), exactly as the GPU path sees device memory.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libttc_oracle.so")

# dtype codes shared with include/ttc_hip.h (ttc_dtype_t)
S, D, C, Z, BF16, F16, E4M3, E5M2 = range(8)

NP_STORAGE = {
    S: np.float32, D: np.float64, C: np.complex64, Z: np.complex128,
    BF16: np.uint16, F16: np.uint16, E4M3: np.uint8, E5M2: np.uint8,
}

_lib = None


def build(force=False):
    """Compile the oracle and the blocked CPU baseline with their Makefile
    (g++, OpenMP)."""
    stale = any(not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(os.path.join(_HERE, src))
                for lib, src in ((_LIB, "ttc_oracle.cpp"), (_BLOCKED_LIB, "ttc_cpu_blocked.cpp")))
    if force or stale:
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB


# ------------------------------------------------ blocked CPU baseline
# ttc_cpu_blocked.cpp: TTC's blocked, vectorised CPU method (W x W in-register
# micro-kernels, AVX2 / AVX-512), a stronger CPU baseline than the scalar
# loop nest; bit-identical to the oracle for s and d.
_BLOCKED_LIB = os.path.join(_HERE, "libttc_cpu_blocked.so")
_blocked = None


def _blocked_lib():
    global _blocked
    if _blocked is None:
        if not os.path.exists(_BLOCKED_LIB):
            build()
        L = ctypes.CDLL(_BLOCKED_LIB)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ttc_cpu_blocked_transpose.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ip, ip, ip, ip,
                                                ctypes.c_int, ctypes.c_int]
        L.ttc_cpu_blocked_transpose.restype = ctypes.c_int
        L.ttc_cpu_blocked_isa.restype = ctypes.c_int
        _blocked = L
    return _blocked


def blocked_isa():
    """2: AVX-512F micro-kernels, 1: AVX2, 0: neither on this host."""
    return _blocked_lib().ttc_cpu_blocked_isa()


def blocked_transpose(t, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None, nthreads=0, isa=0):
    """The blocked CPU baseline, in place on B (t: S or D, A and B of that type)."""
    if beta_is_zero is None:
        beta_is_zero = (beta == 0)
    assert A.flags.c_contiguous and B.flags.c_contiguous
    r = _blocked_lib().ttc_cpu_blocked_transpose(t, A.ctypes.data, B.ctypes.data, float(alpha), float(beta),
                                                 int(bool(beta_is_zero)), len(perm), _iarr(perm), _iarr(size),
                                                 _iarr(lda), _iarr(ldb), int(nthreads), int(isa))
    if r != 0:
        raise ValueError("blocked baseline rejected the arguments (code %d)" % r)
    return B


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ttc_oracle_transpose.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ip, ip, ip, ip, ctypes.c_int]
        L.ttc_oracle_transpose.restype = ctypes.c_int
        L.ttc_oracle_cast_fp8.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_double, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                          ctypes.c_int, ip, ip, ip, ip, ctypes.c_int]
        L.ttc_oracle_cast_fp8.restype = ctypes.c_int
        L.ttc_oracle_convert.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.ttc_oracle_convert.restype = ctypes.c_int
        L.ttc_oracle_trash_cache.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.ttc_oracle_trash_cache.restype = ctypes.c_double
        L.ttc_oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _iarr(v):
    if v is None or len(v) == 0:
        return None
    a = (ctypes.c_int * len(v))(*[int(x) for x in v])
    return a


def transpose(tA, tB, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None, nthreads=0):
    """In place on B (numpy flat buffer).  beta_is_zero defaults to beta == 0,
    the reference's generation-time specialisation (ttc/transpose.py:133-136)."""
    if beta_is_zero is None:
        beta_is_zero = (beta == 0)
    assert A.flags.c_contiguous and B.flags.c_contiguous
    r = lib().ttc_oracle_transpose(tA, tB, A.ctypes.data, B.ctypes.data, float(alpha), float(beta),
                                   int(bool(beta_is_zero)), len(perm), _iarr(perm), _iarr(size),
                                   _iarr(lda), _iarr(ldb), int(nthreads))
    if r != 0:
        raise ValueError("oracle rejected the arguments (code %d)" % r)
    return B


def cast_fp8(tA, tB, A, B, alpha, scale, perm, size, lda=None, ldb=None, nthreads=0):
    amax = ctypes.c_float(0.0)
    r = lib().ttc_oracle_cast_fp8(tA, tB, A.ctypes.data, B.ctypes.data, float(alpha), float(scale),
                                  ctypes.byref(amax), len(perm), _iarr(perm), _iarr(size), _iarr(lda),
                                  _iarr(ldb), int(nthreads))
    if r != 0:
        raise ValueError("oracle rejected the arguments (code %d)" % r)
    return float(amax.value)


def convert(op, x):
    """op 0: f32->bf16 bits, 1: f32->f16 bits, 2: f32->e4m3 (sat), 3: f32->e5m2 (sat)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16 if op < 2 else np.uint8)
    r = lib().ttc_oracle_convert(op, x.ctypes.data, out.ctypes.data, x.size)
    if r != 0:
        raise ValueError("bad op")
    return out


def trash_cache(buf, nthreads=0):
    return lib().ttc_oracle_trash_cache(buf.ctypes.data, buf.size, int(nthreads))


def max_threads():
    return lib().ttc_oracle_max_threads()


def strides(perm, size, lda=None, ldb=None):
    """(sA, sB): element strides of each A-dimension k in A and in B
    (ttc/transpose.py:449-473).  sB[k] is the B stride of B position
    invperm[k]."""
    d = len(size)
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(d)]
    sA = [1] * d
    for k in range(1, d):
        sA[k] = sA[k - 1] * LA[k - 1]
    sBpos = [1] * d
    for q in range(1, d):
        sBpos[q] = sBpos[q - 1] * LB[q - 1]
    sB = [0] * d
    for q in range(d):
        sB[perm[q]] = sBpos[q]
    return sA, sB


def extent_elems(perm, size, lda=None, ldb=None):
    """Number of elements spanned by A and B in memory (with padding)."""
    d = len(size)
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(d)]
    na = int(np.prod(LA[:-1], dtype=np.int64)) * size[-1] if d > 1 else size[0]
    nb = int(np.prod(LB[:-1], dtype=np.int64)) * size[perm[-1]] if d > 1 else size[0]
    return int(na), int(nb)
