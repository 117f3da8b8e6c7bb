"""The reference's own CPU polynomial multiply (modpn), compiled from its sources.

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use this module, as a checker or as
the timed CPU baseline.

``oracle/_ref/libmodpn_fmul.so`` is built by ``oracle/Makefile`` from
modpn/src/FMUL.c, CONSTANTS.c, generalFuncs.c and utilityFunctions.c where
they lie in the reference tree (nothing is copied).  It exists only where the
reference tree was present at build time; every function here returns None
or raises ``RuntimeError`` when it is missing, so callers skip.

Entry points used:
- ``EX_Mont_Mul`` (modpn/src/FMUL.c:2046-2070): plain multiply below degree
  138 (``CutOffPlainFFTMul``, FMUL.h:36), Montgomery FFT multiply above.
- ``EX_Mont_FFTMul_OPT2_AS_GENE`` (FMUL.c:1757-1810): the CPU FFT multiply.
- ``EX_KN_Mont_FFTMul_OPT2_AS_GENE_1`` (FMUL.c:1923-1963): the in-place
  multiply that hands sizes 2^16 and up to ``cumodp_fftmul_uni`` when the
  global ``CUDA_TAG`` is set and ``can_call_fftmul_uni`` agrees, and falls
  back to its CPU FFT when that call fails.  Linked against
  libcumodp_amd.so, it is modpn calling this library through the C ABI.

modpn's Montgomery arithmetic keeps intermediate values in (-p, 2p) as
signed 32-bit integers (utilityFunctions.c:81-106), so it is exact only for
primes below 2^30: for 2013265921 its products are wrong.  The checks here
use 30-bit and smaller Fourier primes.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libmodpn_fmul.so")
REFERENCE = "/root/reference"
_LIB = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class MontP(C.Structure):
    """MONTP_OPT2_AS_GENE (modpn/src/Types.h:427-441): fourteen sfixn fields."""
    _fields_ = [(n, C.c_int) for n in (
        "P", "c", "Npow", "Rpow", "R_Npow", "Base_Rpow", "Base_Npow", "c_sft", "c_pow",
        "N2_Rpow", "Rmask", "R_Nmask", "Max_Root", "R2BRsft")]


def build():
    """Compile _ref/libmodpn_fmul.so when the reference tree is present."""
    import subprocess
    if not os.path.isdir(os.path.join(REFERENCE, "modpn", "src")):
        return False
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, f"REF={REFERENCE}",
                           "_ref/libmodpn_fmul.so"])
    return True


def available():
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        if not available():
            raise RuntimeError(f"{PATH} is not built (needs the reference tree at build time)")
        L = C.CDLL(PATH)
        mp = C.POINTER(MontP)
        s = C.c_int
        L.EX_MontP_Init_OPT2_AS_GENE.argtypes = [mp, s]
        L.EX_Mont_Mul.argtypes = [C.POINTER(s), s, i32p, s, i32p, mp]
        L.EX_Mont_Mul.restype = C.POINTER(s)
        L.EX_Mont_FFTMul_OPT2_AS_GENE.argtypes = [s, s, s, i32p, s, i32p, s, i32p, mp]
        L.EX_Mont_GetNthRoots_OPT2_AS_GENE.argtypes = [s, s, i32p, mp]
        L.EX_KN_Mont_FFTMul_OPT2_AS_GENE_1.argtypes = [s, s, s, i32p, s, i32p, s, i32p, mp]
        L.my_free.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def prime(p):
    pr = MontP()
    lib().EX_MontP_Init_OPT2_AS_GENE(C.byref(pr), int(p))
    return pr


def _flags():
    L = lib()
    return C.c_int.in_dll(L, "Interrupted"), C.c_int.in_dll(L, "PrimeError")


def _check_flags():
    it, pe = _flags()
    if it.value or pe.value:
        it.value = pe.value = 0
        raise ValueError("modpn: transform size beyond the prime's 2-adic order")


def set_cuda_tag(on):
    """Sets modpn's global CUDA_TAG (modpn/src/CONSTANTS.c:55); returns the old value."""
    tag = C.c_int.in_dll(lib(), "CUDA_TAG")
    old = tag.value
    tag.value = 1 if on else 0
    return old


def _log_ceiling(x):
    return max(0, int(x - 1).bit_length())


def mul(A, B, p):
    """EX_Mont_Mul: product of A and B mod p (degree len(A)+len(B)-2)."""
    A = np.ascontiguousarray(A, np.int32)
    B = np.ascontiguousarray(B, np.int32)
    d = C.c_int()
    r = lib().EX_Mont_Mul(C.byref(d), len(A) - 1, A, len(B) - 1, B, C.byref(prime(p)))
    _check_flags()
    out = np.ctypeslib.as_array(r, shape=(d.value + 1,)).copy()
    lib().my_free(C.cast(r, C.c_void_p))
    return out


def fft_mul(A, B, p):
    """EX_Mont_FFTMul_OPT2_AS_GENE: the CPU FFT multiply, whatever the degree."""
    A = np.ascontiguousarray(A, np.int32)
    B = np.ascontiguousarray(B, np.int32)
    dres = len(A) + len(B) - 2
    e = _log_ceiling(dres + 1)
    res = np.zeros(dres + 1, np.int32)
    lib().EX_Mont_FFTMul_OPT2_AS_GENE(1 << e, e, dres, res, len(A) - 1, A, len(B) - 1, B,
                                      C.byref(prime(p)))
    _check_flags()
    return res


def kn_fftmul(A, B, p, cuda_tag):
    """EX_KN_Mont_FFTMul_OPT2_AS_GENE_1 in place on A, as modpn calls it.

    With ``cuda_tag`` set and a transform of 2^16 points or more, modpn calls
    ``cumodp_fftmul_uni(n - 1, A, degA, A, degB, B, p)``, i.e. this library,
    with the output aliasing the first operand.  Returns the n words of A.
    """
    A = np.ascontiguousarray(A, np.int32)
    B = np.ascontiguousarray(B, np.int32)
    da, db = len(A) - 1, len(B) - 1
    e = _log_ceiling(da + db + 1)
    n = 1 << e
    a = np.zeros(n, np.int32)
    a[:da + 1] = A
    b = np.zeros(n, np.int32)
    b[:db + 1] = B
    pr = prime(p)
    roots = np.zeros(n, np.int32)
    L = lib()
    L.EX_Mont_GetNthRoots_OPT2_AS_GENE(e, n, roots, C.byref(pr))
    _check_flags()
    old = set_cuda_tag(cuda_tag)
    try:
        L.EX_KN_Mont_FFTMul_OPT2_AS_GENE_1(n, e, da + db, roots, da, a, db, b, C.byref(pr))
    finally:
        set_cuda_tag(old)
    _check_flags()
    return a
