"""CPU oracle for the NTT polynomial-multiplication path (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline.  The product (``cumodp_amd``) never imports it.

Thin numpy/ctypes wrappers over ``liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``).  Each C function cites the reference file:line whose
semantics it restates; see ``oracle.h``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        s = C.c_int32
        for name in ("orc_add_mod", "orc_sub_mod", "orc_mul_mod"):
            getattr(L, name).argtypes = [s, s, s]
            getattr(L, name).restype = s
        L.orc_neg_mod.argtypes = [s, s]; L.orc_neg_mod.restype = s
        L.orc_inv_mod.argtypes = [s, s]; L.orc_inv_mod.restype = s
        L.orc_pow_mod.argtypes = [s, s, s]; L.orc_pow_mod.restype = s
        L.orc_fourier_degree.argtypes = [s]; L.orc_fourier_degree.restype = s
        L.orc_primitive_root.argtypes = [s, s]; L.orc_primitive_root.restype = s
        L.orc_get_powers.argtypes = [s, i32p, s, s]
        L.orc_dft_naive.argtypes = [s, i32p, i32p, s, s]
        L.orc_gsnn_fft.argtypes = [s, s, i32p, s, s]
        L.orc_stockham.argtypes = [i32p, s, s, s, s]
        L.orc_inverse_stockham.argtypes = [i32p, s, s, s, s]
        L.orc_list_stockham.argtypes = [i32p, s, s, s, s]
        L.orc_plain_polymul.argtypes = [s, i32p, s, i32p, s, i32p, s]
        L.orc_ntt_polymul.argtypes = [i32p, s, i32p, s, i32p, s]
        L.orc_polymul_items.argtypes = [C.c_int64, i64p, i32p, i32p, i32p, s, s, s]
        L.orc_stockham_poly_mul_host.argtypes = [s, i32p, s, i32p, s, i32p, s]
        L.orc_stockham_poly_mul_host.restype = C.c_int
        L.orc_zmul_i64_schoolbook.argtypes = [i64p, C.c_int64, i64p, C.c_int64, u64p]
        L.orc_zmul_mw_schoolbook.argtypes = [u64p, C.c_int64, C.c_int, u64p, C.c_int64,
                                             C.c_int, u64p, C.c_int]
        L.orc_bit_width_mw.argtypes = [u64p, C.c_int64, C.c_int]
        L.orc_bit_width_mw.restype = C.c_int
        L.orc_limb_split.argtypes = [u64p, C.c_int64, C.c_int, C.c_int, C.c_int, s, i32p]
        L.orc_scale_x.argtypes = [i32p, C.c_int64, C.c_int, s, s]
        L.orc_two_convolution_zmul.argtypes = [u64p, C.c_int64, C.c_int, u64p, C.c_int64,
                                               C.c_int, u64p, C.c_int]
        L.orc_two_convolution_zmul.restype = C.c_int
        L.orc_subproduct_level.argtypes = [i32p, C.c_int64, C.c_int, s, i32p]
        L.orc_poly_eval.argtypes = [i32p, C.c_int64, i32p, C.c_int64, s, i32p]
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_splitmix64.restype = C.c_uint64
        _LIB = L
    return _LIB


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- scalar arithmetic ------------------------------------------------------
def add_mod(a, b, p): return lib().orc_add_mod(a, b, p)
def sub_mod(a, b, p): return lib().orc_sub_mod(a, b, p)
def neg_mod(a, p): return lib().orc_neg_mod(a, p)
def mul_mod(a, b, p): return lib().orc_mul_mod(a, b, p)
def inv_mod(a, p): return lib().orc_inv_mod(a, p)
def pow_mod(a, e, p): return lib().orc_pow_mod(a, e, p)
def fourier_degree(p): return lib().orc_fourier_degree(p)
def primitive_root(e, p): return lib().orc_primitive_root(e, p)


# ---- transforms -------------------------------------------------------------
def get_powers(n, w, p):
    W = np.zeros(n, np.int32)
    lib().orc_get_powers(n, W, w, p)
    return W


def dft_naive(X, w, p):
    X = _i32(X)
    Y = np.zeros_like(X)
    lib().orc_dft_naive(len(X), X, Y, w, p)
    return Y


def gsnn_fft(X, w, p):
    A = _i32(X).copy()
    n = len(A)
    lib().orc_gsnn_fft(n, n.bit_length() - 1, A, p, w)
    return A


def stockham(X, w, p):
    A = _i32(X).copy()
    n = len(A)
    lib().orc_stockham(A, n, n.bit_length() - 1, w, p)
    return A


def inverse_stockham(X, w, p):
    A = _i32(X).copy()
    n = len(A)
    lib().orc_inverse_stockham(A, n, n.bit_length() - 1, w, p)
    return A


def list_stockham(X, m, k, w, p):
    A = _i32(X).copy()
    lib().orc_list_stockham(A, m, k, w, p)
    return A


# ---- multiplication mod p ---------------------------------------------------
def plain_polymul(F, G, p):
    F, G = _i32(F), _i32(G)
    H = np.zeros(len(F) + len(G) - 1, np.int32)
    lib().orc_plain_polymul(len(H) - 1, H, len(F) - 1, F, len(G) - 1, G, p)
    return H


def ntt_polymul(F, G, p):
    F, G = _i32(F), _i32(G)
    H = np.zeros(len(F) + len(G) - 1, np.int32)
    lib().orc_ntt_polymul(H, len(F) - 1, F, len(G) - 1, G, p)
    return H


def poly_mul(F, G, p):
    """Exact product mod p: schoolbook when small or when p has no root of
    unity of the transform's order, CPU NTT otherwise."""
    e = max(0, (len(F) + len(G) - 2).bit_length())
    if (min(len(F), len(G)) <= 64 or len(F) * len(G) <= 1 << 16
            or e > fourier_degree(p)):
        return plain_polymul(F, G, p)
    return ntt_polymul(F, G, p)


def polymul_items(desc, F, G, nh, p, plain_cut=64, threads=1):
    """Independent products (desc: (count, 5) int64 rows of f_off, g_off,
    h_off, df, dg) on `threads` OpenMP threads; returns the nh-word H."""
    desc = np.ascontiguousarray(desc, np.int64)
    H = np.zeros(nh, np.int32)
    lib().orc_polymul_items(desc.shape[0], desc.ravel(), _i32(F), _i32(G), H, p, plain_cut,
                            threads)
    return H


def stockham_poly_mul_host(dh, H, F, G, p):
    """Host-wrapper semantics of src/stockham.cu:562-611 (overrun fixed)."""
    F, G = _i32(F), _i32(G)
    return lib().orc_stockham_poly_mul_host(dh, H, len(F) - 1, F, len(G) - 1, G, p)


# ---- exact integer multiplication -------------------------------------------
def words_to_ints(w):
    """(n, W) uint64 two's-complement words -> list of Python ints."""
    w = np.asarray(w, dtype=np.uint64)
    n, W = w.shape
    out = []
    for i in range(n):
        v = 0
        for j in reversed(range(W)):
            v = (v << 64) | int(w[i, j])
        if v >> (64 * W - 1):
            v -= 1 << (64 * W)
        out.append(v)
    return out


def ints_to_words(vals, W):
    out = np.zeros((len(vals), W), np.uint64)
    mask = (1 << 64) - 1
    for i, v in enumerate(vals):
        v &= (1 << (64 * W)) - 1
        for j in range(W):
            out[i, j] = (v >> (64 * j)) & mask
    return out


def zmul_i64(a, b):
    """Exact product of int64 polynomials, as (n, 3) 192-bit words."""
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    c = np.zeros((len(a) + len(b) - 1) * 3, np.uint64)
    lib().orc_zmul_i64_schoolbook(a, len(a), b, len(b), c)
    return c.reshape(-1, 3)


def zmul_mw(a, b, wc):
    """Schoolbook on (n, W) two's-complement words; returns (n, wc)."""
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    na, wa = a.shape
    nb, wb = b.shape
    c = np.zeros((na + nb - 1) * wc, np.uint64)
    lib().orc_zmul_mw_schoolbook(a.ravel(), na, wa, b.ravel(), nb, wb, c, wc)
    return c.reshape(-1, wc)


def bit_width(a):
    a = np.ascontiguousarray(a, np.uint64)
    return lib().orc_bit_width_mw(a.ravel(), a.shape[0], a.shape[1])


def limb_split(a, K, M, prime):
    a = np.ascontiguousarray(a, np.uint64)
    out = np.zeros(a.shape[0] * K, np.int32)
    lib().orc_limb_split(a.ravel(), a.shape[0], a.shape[1], K, M, prime, out)
    return out


def scale_x(A, K, theta, prime):
    A = _i32(A).copy()
    lib().orc_scale_x(A, len(A), K, theta, prime)
    return A


def two_convolution_zmul(a, b, wc):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    na, wa = a.shape
    nb, wb = b.shape
    c = np.zeros((na + nb - 1) * wc, np.uint64)
    rc = lib().orc_two_convolution_zmul(a.ravel(), na, wa, b.ravel(), nb, wb, c, wc)
    if rc != 0:
        raise ValueError("no two-convolution base satisfies the CRT bound")
    return c.reshape(-1, wc)


# ---- subproduct tree --------------------------------------------------------
def subproduct_level(points, lev, p):
    pts = _i32(points)
    n = len(pts)
    out = np.zeros((n >> lev) * ((1 << lev) + 1), np.int32)
    lib().orc_subproduct_level(pts, n, lev, p, out)
    return out.reshape(n >> lev, (1 << lev) + 1)


# ---- multipoint evaluation ---------------------------------------------------
def poly_eval(F, X, p):
    """F(X[j]) mod p by Horner's rule."""
    F, X = _i32(F), _i32(X)
    out = np.zeros(len(X), np.int32)
    lib().orc_poly_eval(F, len(F), X, len(X), p, out)
    return out
