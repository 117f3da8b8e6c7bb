"""cumodp_amd -- MI355X (gfx950) NTT polynomial multiplication over 32-bit
Fourier primes, a drop-in for CUMODP's multiply path.

The product is ``libcumodp_amd.so`` (hand-written HIP kernels behind a C ABI,
declared in ``include/*.h``).  This package is a thin ctypes binding used by
the tests and ``bench.py``; it mirrors the reference's entry points with the
same names, argument meaning and error behaviour:

* ``stockham_poly_mul_host(dh, H, df, F, dg, G, p)``   src/stockham.cu:562
* ``cumodp_fftmul_uni(dh, H, df, F, dg, G, p)``        src/modpn.cu:250
* ``cumodp_fft_uni / cumodp_invfft_uni``               src/modpn.cu:182,199
* ``stockham_host / inverse_stockham_host``            src/stockham.cu:479,490
* ``list_stockham_host``                               src/list_stockham.cu:656

There is no CPU fallback: importing works without a GPU (for symbol checks),
but every compute call goes to the HIP library and fails loudly if the library
is missing or the device call fails.
"""
import ctypes as C
import os

import numpy as np

from .workload import *  # noqa: F401,F403  (seeded generators shared by tests and bench)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcumodp_amd.so")
_LIB = None

CUMODP_SUCCESS = 0x000
CUMODP_FAILURE = 0x001
CUMODP_ASSUMPTION_ERROR = 0x002
CUMODP_FFT_ERROR = 0x200
CUMODP_FOURIER_DEGREE_TOO_SMALL = 0x201
CUMODP_FFT_SIZE_TOO_LARGE = 0x202
CUMODP_FFT_SIZE_TOO_SMALL = 0x203
CUMODP_OUT_OF_MEMORY = 0x401

# mangled names of the C++-linkage drop-ins (same signatures as the reference)
CXX_SYMBOLS = {
    "stockham_dev_w": "_Z12stockham_devPiiiii",
    "stockham_dev_W": "_Z12stockham_devPiiiPKii",
    "inverse_stockham_dev_w": "_Z20inverse_stockham_devPiiiii",
    "inverse_stockham_dev_W": "_Z20inverse_stockham_devPiiiPKii",
    "stockham_host": "_Z13stockham_hostPiiiii",
    "inverse_stockham_host": "_Z21inverse_stockham_hostPiiiii",
    "stockham_poly_mul_dev": "_Z21stockham_poly_mul_deviiPiS_i",
    "stockham_poly_mul_host": "_Z22stockham_poly_mul_hostiPiiPKiiS1_i",
    "list_stockham_host": "_Z18list_stockham_hostPiiiii",
    "list_stockham_dev_w": "_Z17list_stockham_devPiiiii",
    "list_stockham_dev_W": "_Z17list_stockham_devPiiiPKii",
    "list_inv_stockham_host": "_Z22list_inv_stockham_hostPiiiii",
    "get_powers_binary": "_Z17get_powers_binaryiPiii",
    "pointwise_mul_dev": "_Z17pointwise_mul_deviiPiPKii",
    "scale_vector_dev": "_Z16scale_vector_deviiPii",
    "expand_to_fft_dev": "_Z17expand_to_fft_deviiPi",
    "reset_vector_dev": "_Z16reset_vector_deviPi",
    "bivariate_stockham_dev": "_Z22bivariate_stockham_devPiiPKiiS1_i",
    "inverse_bivariate_stockham_dev": "_Z30inverse_bivariate_stockham_devPiiPKiiS1_i",
    "bivariate_stockham_host": "_Z23bivariate_stockham_hostPiiiiii",
    "inverse_bivariate_stockham_host": "_Z31inverse_bivariate_stockham_hostPiiiiii",
    "bi_stockham_poly_mul_dev": "_Z24bi_stockham_poly_mul_deviiiiPiS_i",
    "list_bivariate_stockham_dev_W": "_Z27list_bivariate_stockham_devPiiiPKiiS1_i",
    "list_bivariate_stockham_dev_w": "_Z27list_bivariate_stockham_devPiiiiiii",
    "mulCUDA": "_Z7mulCUDAPiS_S_iiii",
    "fastEvaluation": "_Z14fastEvaluationPiS_iii",
    "onlySubproductTree": "_Z18onlySubproductTreePiS_iii",
    "onlyTrees": "_Z9onlyTreesPiiii",
}

MAX_LEVEL = 27


class PolyEvalSteps(C.Structure):
    """struct PolyEvalSteps of include/fastPolyEvaluation.h (device pointers)"""
    _fields_ = [(f, C.c_void_p * MAX_LEVEL) for f in ("Ml", "Mr", "InvMl", "InvMr", "fL", "fR")] + \
               [("timeSubTree", C.c_float), ("timeSinvTree", C.c_float), ("timeEva", C.c_float)]


class CumodpError(RuntimeError):
    def __init__(self, fn, code):
        msg = ""
        if _LIB is not None:
            msg = _LIB.cumodp_last_error().decode(errors="replace")
        super().__init__(f"{fn} failed with status 0x{code:03x}: {msg}")
        self.code = code


class MulItem(C.Structure):
    """cumodp_mul_item of include/cumodp_amd.h"""
    _fields_ = [("f_off", C.c_int64), ("g_off", C.c_int64), ("h_off", C.c_int64),
                ("df", C.c_int32), ("dg", C.c_int32)]


class ZmulItem(C.Structure):
    """cumodp_zmul_item of include/cumodp_amd.h (word offsets and widths)"""
    _fields_ = [("a_off", C.c_int64), ("b_off", C.c_int64), ("c_off", C.c_int64),
                ("na", C.c_int64), ("nb", C.c_int64), ("wa", C.c_int32), ("wb", C.c_int32),
                ("wc", C.c_int32), ("bits", C.c_int32)]


def lib():
    """Load libcumodp_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                               "python -m cumodp_amd.build")
        L = C.CDLL(LIB_PATH)
        s, vp, sp = C.c_int, C.c_void_p, C.c_void_p
        err = C.c_int
        L.cumodp_fftmul_uni.argtypes = [s, sp, s, sp, s, sp, s]
        L.cumodp_fftmul_uni.restype = err
        L.cumodp_fft_uni.argtypes = [sp, s, s, s, s]
        L.cumodp_fft_uni.restype = err
        L.cumodp_invfft_uni.argtypes = [sp, s, s, s, s]
        L.cumodp_invfft_uni.restype = err
        L.can_call_fftmul_uni.argtypes = [s, s]
        L.can_call_fftmul_uni.restype = s
        L.global_memory_in_bytes.restype = C.c_uint
        L.global_memory_in_megabytes.restype = C.c_float
        L.cumodp_dev_alloc.argtypes = [C.c_size_t]
        L.cumodp_dev_alloc.restype = vp
        L.cumodp_dev_free.argtypes = [vp]
        L.cumodp_dev_free.restype = err
        for n in ("cumodp_memcpy_h2d", "cumodp_memcpy_d2h", "cumodp_memcpy_d2d"):
            getattr(L, n).argtypes = [vp, vp, C.c_size_t]
            getattr(L, n).restype = err
        L.cumodp_synchronize.restype = err
        L.cumodp_stream_create.restype = vp
        L.cumodp_stream_destroy.argtypes = [vp]
        L.cumodp_stream_destroy.restype = err
        L.cumodp_stream_synchronize.argtypes = [vp]
        L.cumodp_stream_synchronize.restype = err
        L.cumodp_last_error.restype = C.c_char_p
        L.cumodp_fftmul_uni_dev.argtypes = [s, vp, s, vp, s, vp, s, vp]
        L.cumodp_fftmul_uni_dev.restype = err
        L.cumodp_fftmul_uni_dev_noalloc.argtypes = [s, vp, s, vp, s, vp, s, vp]
        L.cumodp_fftmul_uni_dev_noalloc.restype = err
        L.cumodp_list_fftmul_dev.argtypes = [C.POINTER(MulItem), C.c_int64, vp, vp, vp, s, vp]
        L.cumodp_list_fftmul_dev.restype = err
        L.cumodp_list_plainmul_dev.argtypes = [C.POINTER(MulItem), C.c_int64, vp, vp, vp, s, vp]
        L.cumodp_list_plainmul_dev.restype = err
        L.cumodp_plain_mul_dev.argtypes = [vp, vp, s, vp, s, s, vp]
        L.cumodp_plain_mul_dev.restype = err
        L.cumodp_list_fftmul_dev_percall.argtypes = [C.POINTER(MulItem), C.c_int64, vp, vp, vp,
                                                     s, vp]
        L.cumodp_list_fftmul_dev_percall.restype = err
        L.cumodp_fft_bivariate.argtypes = [sp, s, s, s, s, s]
        L.cumodp_fft_bivariate.restype = err
        L.cumodp_invfft_bivariate.argtypes = [sp, s, s, s, s, s]
        L.cumodp_invfft_bivariate.restype = err
        L.cumodp_profile_enable.argtypes = [s]
        L.cumodp_profile_enable.restype = err
        L.cumodp_profile_count.restype = s
        L.cumodp_profile_name.argtypes = [s]
        L.cumodp_profile_name.restype = C.c_char_p
        L.cumodp_profile_read.argtypes = [s, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.cumodp_profile_read.restype = err
        L.cumodp_tune_set_split.argtypes = [s, s]
        L.cumodp_tune_set_split.restype = err
        L.cumodp_tune_get_split.argtypes = [s]
        L.cumodp_tune_get_split.restype = s
        L.cumodp_tune_set_list_tile.argtypes = [s]
        L.cumodp_tune_set_list_tile.restype = err
        L.cumodp_tune_get_list_tile.restype = s
        L.cumodp_tune_set_list_plain.argtypes = [s]
        L.cumodp_tune_set_list_plain.restype = err
        L.cumodp_tune_get_list_plain.restype = s
        L.cumodp_tune_set_list_persistent.argtypes = [s]
        L.cumodp_tune_set_list_persistent.restype = err
        L.cumodp_tune_get_list_persistent.restype = s
        L.cumodp_tune_set_zmul_merge.argtypes = [s]
        L.cumodp_tune_set_zmul_merge.restype = err
        L.cumodp_tune_get_zmul_merge.restype = s
        L.cumodp_tune_set_col_tile.argtypes = [s]
        L.cumodp_tune_set_col_tile.restype = err
        L.cumodp_tune_get_col_tile.restype = s
        i64 = C.c_int64
        L.cumodp_two_convolution_poly_mul.argtypes = [vp, i64, s, vp, i64, s, vp, s]
        L.cumodp_two_convolution_poly_mul.restype = err
        L.cumodp_two_convolution_poly_mul_dev.argtypes = [vp, i64, s, vp, i64, s, vp, s, s, vp]
        L.cumodp_two_convolution_poly_mul_dev.restype = err
        L.cumodp_find_largest_bit_width.argtypes = [vp, i64, s, vp, i64, s]
        L.cumodp_find_largest_bit_width.restype = s
        L.cumodp_convert_to_modular_bivariate.argtypes = [vp, i64, s, s, s, s, vp]
        L.cumodp_convert_to_modular_bivariate.restype = err
        L.cumodp_scale_x_argument_dev.argtypes = [vp, i64, s, s, s, vp]
        L.cumodp_scale_x_argument_dev.restype = err
        L.cumodp_two_convolution_plan.argtypes = [s, i64, i64, C.POINTER(s), C.POINTER(s),
                                                  C.POINTER(s)]
        L.cumodp_two_convolution_plan.restype = s
        L.cumodp_subproduct_tree_words.argtypes = [s]
        L.cumodp_subproduct_tree_words.restype = i64
        L.cumodp_subproduct_tree_offset.argtypes = [s, s]
        L.cumodp_subproduct_tree_offset.restype = i64
        L.cumodp_subproduct_tree_dev.argtypes = [vp, s, s, vp, vp]
        L.cumodp_subproduct_tree_dev.restype = err
        L.cumodp_subproduct_tree.argtypes = [vp, s, s, vp]
        L.cumodp_subproduct_tree.restype = err
        L.cumodp_fast_evaluation.argtypes = [vp, vp, s, s, vp]
        L.cumodp_fast_evaluation.restype = err
        L.cumodp_fast_evaluation_dev.argtypes = [vp, vp, s, s, vp, vp]
        L.cumodp_fast_evaluation_dev.restype = err
        L.cumodp_two_convolution_list_dev.argtypes = [C.POINTER(ZmulItem), i64, vp, vp, vp, vp]
        L.cumodp_two_convolution_list_dev.restype = err
        L.cumodp_host_alloc.argtypes = [C.c_size_t]
        L.cumodp_host_alloc.restype = vp
        L.cumodp_host_free.argtypes = [vp]
        L.cumodp_host_free.restype = err
        for n in ("cumodp_memcpy_h2d_async", "cumodp_memcpy_d2h_async"):
            getattr(L, n).argtypes = [vp, vp, C.c_size_t, vp]
            getattr(L, n).restype = err
        L.cumodp_stream_query.argtypes = [vp]
        L.cumodp_stream_query.restype = s
        L.cumodp_graph_begin.argtypes = [vp]
        L.cumodp_graph_begin.restype = err
        L.cumodp_graph_end.argtypes = [vp]
        L.cumodp_graph_end.restype = vp
        L.cumodp_graph_launch.argtypes = [vp, vp]
        L.cumodp_graph_launch.restype = err
        L.cumodp_graph_destroy.argtypes = [vp]
        L.cumodp_graph_destroy.restype = err
        _LIB = L
    return _LIB


def cxx(name):
    """ctypes handle of a C++-linkage drop-in by its reference name."""
    return getattr(lib(), CXX_SYMBOLS[name])


def _check(fn, code):
    if code != CUMODP_SUCCESS:
        raise CumodpError(fn, code)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- device buffers ----------------------------------------------------------
class DeviceArray:
    """A device allocation owned by the library's allocator (hipMalloc)."""

    def __init__(self, nwords, dtype=np.int32):
        self.dtype = np.dtype(dtype)
        self.n = int(nwords)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = lib().cumodp_dev_alloc(max(self.nbytes, 4))
        if not self.ptr:
            raise CumodpError("cumodp_dev_alloc", CUMODP_OUT_OF_MEMORY)

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        _check("cumodp_memcpy_h2d", lib().cumodp_memcpy_h2d(
            C.c_void_p(self.ptr + offset * self.dtype.itemsize), _ptr(a), a.nbytes))

    def download(self, count=None, offset=0):
        count = self.n - offset if count is None else count
        out = np.empty(count, self.dtype)
        _check("cumodp_memcpy_d2h", lib().cumodp_memcpy_d2h(
            _ptr(out), C.c_void_p(self.ptr + offset * self.dtype.itemsize), out.nbytes))
        return out

    def addr(self, offset=0):
        return C.c_void_p(self.ptr + offset * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            lib().cumodp_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check("cumodp_synchronize", lib().cumodp_synchronize())


# ---- reference-named entry points (host arrays) -------------------------------
def stockham_poly_mul_host(dh, H, df, F, dg, G, p):
    """src/stockham.cu:562: writes min(dh+1, n) words of H in place."""
    assert H.dtype == np.int32 and H.flags.c_contiguous
    F, G = _i32(F), _i32(G)
    cxx("stockham_poly_mul_host")(C.c_int(dh), _ptr(H), C.c_int(df), _ptr(F), C.c_int(dg),
                                  _ptr(G), C.c_int(p))


def cumodp_fftmul_uni(dh, H, df, F, dg, G, p):
    """src/modpn.cu:250: returns the cumodp_err status (0 on success)."""
    assert H.dtype == np.int32 and H.flags.c_contiguous
    F, G = _i32(F), _i32(G)
    return lib().cumodp_fftmul_uni(dh, _ptr(H), df, _ptr(F), dg, _ptr(G), p)


def poly_mul(F, G, p):
    """Product F*G mod p (all df+dg+1 coefficients) through cumodp_fftmul_uni."""
    F, G = _i32(F), _i32(G)
    H = np.zeros(len(F) + len(G) - 1, np.int32)
    _check("cumodp_fftmul_uni", cumodp_fftmul_uni(len(H) - 1, H, len(F) - 1, F, len(G) - 1, G, p))
    return H


def fft_uni(X, w, p):
    """Natural-order DFT (cumodp_fft_uni); returns a new array."""
    X = _i32(X).copy()
    n = len(X)
    _check("cumodp_fft_uni", lib().cumodp_fft_uni(_ptr(X), n, n.bit_length() - 1, w, p))
    return X


def invfft_uni(X, w, p):
    X = _i32(X).copy()
    n = len(X)
    _check("cumodp_invfft_uni", lib().cumodp_invfft_uni(_ptr(X), n, n.bit_length() - 1, w, p))
    return X


def list_stockham_host(X, m, k, w, p):
    """I_m (x) DFT_{2^k} of a host array (in a copy)."""
    X = _i32(X).copy()
    cxx("list_stockham_host")(_ptr(X), C.c_int(m), C.c_int(k), C.c_int(w), C.c_int(p))
    return X


def fftmul_uni_dev(dh, Hd, df, Fd, dg, Gd, p, stream=None, noalloc=False):
    """cumodp_fftmul_uni_dev on DeviceArray/int addresses (stream ordered);
    noalloc=True calls cumodp_fftmul_uni_dev_noalloc (no device allocation,
    transform size at most 2^22, tables built by an earlier allocating call)."""
    a = lambda x: x.addr() if isinstance(x, DeviceArray) else C.c_void_p(x)
    fn = "cumodp_fftmul_uni_dev_noalloc" if noalloc else "cumodp_fftmul_uni_dev"
    _check(fn, getattr(lib(), fn)(dh, a(Hd), df, a(Fd), dg, a(Gd), p, C.c_void_p(stream)))


def list_fftmul_dev(items, Fd, Gd, Hd, p, stream=None, per_call=False):
    """cumodp_list_fftmul_dev (or the per-call form): items is a ctypes array
    of MulItem."""
    fn = "cumodp_list_fftmul_dev_percall" if per_call else "cumodp_list_fftmul_dev"
    _check(fn, getattr(lib(), fn)(items, len(items), Fd.addr(), Gd.addr(), Hd.addr(), p,
                                  C.c_void_p(stream)))


def list_plainmul_dev(items, Fd, Gd, Hd, p, stream=None):
    """cumodp_list_plainmul_dev: schoolbook products of the same items."""
    _check("cumodp_list_plainmul_dev", lib().cumodp_list_plainmul_dev(
        items, len(items), Fd.addr(), Gd.addr(), Hd.addr(), p, C.c_void_p(stream)))


def plain_mul_dev(Cd, Ad, n, Bd, m, p, stream=None):
    """cumodp_plain_mul_dev on device addresses (DeviceArray or int)."""
    def a(x):
        v = x.addr() if hasattr(x, "addr") else x
        return v if isinstance(v, C.c_void_p) else C.c_void_p(v)
    _check("cumodp_plain_mul_dev", lib().cumodp_plain_mul_dev(
        a(Cd), a(Ad), n, a(Bd), m, p, C.c_void_p(stream)))


def mulCUDA(A, B, p, C_out=None):
    """The reference's plain multiply mulCUDA(A, B, C, n, m, Prime, verify)
    (src/opt_plain_mul.cu:258): returns C = A * B mod p (len(A)+len(B)-1
    words).  C_out, when given, is the int32 output array (left untouched on
    failure)."""
    A, B = _i32(A), _i32(B)
    out = np.zeros(len(A) + len(B) - 1, np.int32) if C_out is None else C_out
    fn = cxx("mulCUDA")
    fn.restype = None
    fn(_ptr(A), _ptr(B), _ptr(out), len(A), len(B), p, 1)
    return out


def make_items(df, dg):
    """Packed layout for pairs of degrees df[i], dg[i]: ctypes MulItem array
    and the total F, G, H word counts."""
    n = len(df)
    items = (MulItem * n)()
    fo = go = ho = 0
    for i in range(n):
        a, b = int(df[i]), int(dg[i])
        items[i].f_off, items[i].g_off, items[i].h_off = fo, go, ho
        items[i].df, items[i].dg = a, b
        fo += a + 1
        go += b + 1
        ho += a + b + 1
    return items, fo, go, ho


# ---- optional per-kernel timing ----------------------------------------------
def profile_enable(on=True):
    lib().cumodp_profile_enable(1 if on else 0)


def profile_reset():
    lib().cumodp_profile_reset()


def profile_read():
    """{kernel name: (total_ms, launches)} for kernels launched while enabled."""
    L = lib()
    out = {}
    for i in range(L.cumodp_profile_count()):
        t, c = C.c_double(), C.c_int64()
        _check("cumodp_profile_read", L.cumodp_profile_read(i, C.byref(t), C.byref(c)))
        if c.value:
            out[L.cumodp_profile_name(i).decode()] = (t.value, c.value)
    return out


# ---- exact integer multiplication (two convolutions + CRT) -------------------
def _words(a, W=None):
    """int64 array or (n, W) uint64 words -> contiguous (n, W) uint64."""
    a = np.asarray(a)
    if a.ndim == 1:
        return np.ascontiguousarray(a.astype(np.int64).view(np.uint64).reshape(-1, 1))
    return np.ascontiguousarray(a, dtype=np.uint64)


def two_convolution_poly_mul(a, b, wc):
    """Exact product of integer polynomials (int64 arrays or (n, W) uint64
    two's-complement words); returns (na+nb-1, wc) uint64 words."""
    A, B = _words(a), _words(b)
    na, wa = A.shape
    nb, wb = B.shape
    out = np.zeros((na + nb - 1, wc), np.uint64)
    _check("cumodp_two_convolution_poly_mul", lib().cumodp_two_convolution_poly_mul(
        _ptr(A), na, wa, _ptr(B), nb, wb, _ptr(out), wc))
    return out


def two_convolution_plan(bits, na, nb):
    npr, K, M = C.c_int(), C.c_int(), C.c_int()
    if not lib().cumodp_two_convolution_plan(bits, na, nb, C.byref(npr), C.byref(K), C.byref(M)):
        return None
    return npr.value, K.value, M.value


# ---- subproduct tree ----------------------------------------------------------
def subproduct_tree(points, p):
    """All levels of the subproduct tree over len(points) = 2^k points mod p:
    a list whose entry i is a (2^(k-i), 2^i + 1) array of monic products."""
    pts = _i32(points)
    k = len(pts).bit_length() - 1
    assert len(pts) == 1 << k
    L = lib()
    tree = np.zeros(L.cumodp_subproduct_tree_words(k), np.int32)
    _check("cumodp_subproduct_tree", L.cumodp_subproduct_tree(_ptr(pts), k, p, _ptr(tree)))
    out = []
    for i in range(k + 1):
        off = L.cumodp_subproduct_tree_offset(k, i)
        out.append(tree[off:off + (1 << k)  + (1 << (k - i))].reshape(1 << (k - i), (1 << i) + 1))
    return out


def fast_evaluation(F, points, p):
    """F(points[j]) mod p for len(points) = 2^k points, F of degree < 2^k."""
    pts = _i32(points)
    k = len(pts).bit_length() - 1
    assert len(pts) == 1 << k
    Fw = np.zeros(1 << k, np.int32)
    Fw[:len(F)] = _i32(F)
    out = np.zeros(1 << k, np.int32)
    _check("cumodp_fast_evaluation", lib().cumodp_fast_evaluation(_ptr(Fw), _ptr(pts), k, p,
                                                                  _ptr(out)))
    return out


def eval_steps(fn, points, p, F=None, flag=0):
    """Calls the reference-named driver fn ("fastEvaluation", "onlySubproductTree"
    or "onlyTrees") and returns its PolyEvalSteps as host arrays: a dict
    field -> {level: int32 array}, every device array freed afterwards.
    None when the call failed (all pointers null)."""
    pts = _i32(points)
    k = len(pts).bit_length() - 1
    assert len(pts) == 1 << k
    f = cxx(fn)
    f.restype = PolyEvalSteps
    if fn == "onlyTrees":
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        r = f(_ptr(pts), k, p, flag)
    else:
        Fw = np.zeros(1 << k, np.int32)
        if F is not None:
            Fw[:len(F)] = _i32(F)
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        r = f(_ptr(Fw), _ptr(pts), k, p, flag)
    nh = (1 << k) // 2
    out, any_set = {}, False
    for name in ("Ml", "Mr", "InvMl", "InvMr", "fL", "fR"):
        arr = getattr(r, name)
        levels = {}
        for i in range(MAX_LEVEL):
            if not arr[i]:
                continue
            any_set = True
            if name in ("Ml", "Mr"):
                words = (nh >> i) * ((1 << i) + 1 if i < 8 else 1 << i)
            else:
                words = nh
            h = np.empty(words, np.int32)
            _check("cumodp_memcpy_d2h", lib().cumodp_memcpy_d2h(_ptr(h), C.c_void_p(arr[i]),
                                                               h.nbytes))
            lib().cumodp_dev_free(C.c_void_p(arr[i]))
            levels[i] = h
        out[name] = levels
    return out if any_set else None


# ---- graphs of repeated shapes -------------------------------------------------
class Graph:
    """A captured sequence of stream-ordered calls, replayed with one launch.

        g = Graph.capture(stream, fn)   # fn() issues cumodp_*_dev calls on stream
        g.launch(stream)

    fn runs once eagerly first (scratch and twiddle tables are allocated
    then), and again under capture."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def capture(cls, stream, fn, warm=True):
        L = lib()
        if warm:
            fn()
            _check("cumodp_stream_synchronize", L.cumodp_stream_synchronize(C.c_void_p(stream)))
        _check("cumodp_graph_begin", L.cumodp_graph_begin(C.c_void_p(stream)))
        try:
            fn()
        finally:
            h = L.cumodp_graph_end(C.c_void_p(stream))
        if not h:
            raise CumodpError("cumodp_graph_end", CUMODP_FAILURE)
        return cls(h)

    def launch(self, stream):
        _check("cumodp_graph_launch", lib().cumodp_graph_launch(C.c_void_p(self.handle),
                                                                C.c_void_p(stream)))

    def destroy(self):
        if self.handle:
            lib().cumodp_graph_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
