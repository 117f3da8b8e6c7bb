"""The C-ABI boundary, checked without a GPU.

* libcumodp_amd.so loads and exports every function include/*.h declares
  (C names directly, C++ drop-ins under the reference's mangled signatures);
* the host-side argument checks return the documented status codes before
  any device call is made.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    """(header, name, is_c) for every CUMODP_API declaration."""
    out = []
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        is_c = 'extern "C"' in text
        for m in re.finditer(r"^CUMODP_API\s+[^;(]*?\**\s*(\w+)\s*\(", text, re.M):
            out.append((fn, m.group(1), is_c))
    return out


def exported(path):
    r = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                       check=True)
    raw = {l.split()[-1] for l in r.stdout.splitlines() if l.strip()}
    dem = subprocess.run(["c++filt"], input="\n".join(sorted(raw)), capture_output=True,
                         text=True, check=True).stdout.splitlines()
    names = set()
    for d in dem:
        names.add(d.split("(")[0].strip())
    return raw, names


def test_headers_declare_the_path():
    names = {n for _, n, _ in declared_functions()}
    for required in ("stockham_poly_mul_host", "stockham_poly_mul_dev", "cumodp_fftmul_uni",
                     "list_stockham_dev", "stockham_dev", "cumodp_fft_uni",
                     "cumodp_fftmul_uni_dev"):
        assert required in names


def test_library_exports_every_declared_symbol(cm):
    raw, names = exported(cm.LIB_PATH)
    missing = [(h, n) for h, n, is_c in declared_functions()
               if (n not in raw if is_c else n not in names)]
    assert not missing, f"declared but not exported: {missing}"
    # the C++ drop-ins carry the reference's exact mangled signatures
    for mangled in cm.CXX_SYMBOLS.values():
        assert mangled in raw, mangled
        assert hasattr(cm.lib(), mangled)


def test_fftmul_rejects_small_sizes_before_device(cm):
    F = np.arange(10, dtype=np.int32)
    H = np.full(19, -1, np.int32)
    rc = cm.cumodp_fftmul_uni(18, H, 9, F, 9, F, cm.P1)
    assert rc == cm.CUMODP_FFT_SIZE_TOO_SMALL
    assert (H == -1).all()


def test_fftmul_rejects_bad_primes_before_device(cm):
    F = np.ones(300, np.int32)
    H = np.full(599, -1, np.int32)
    # 1000 is even
    assert cm.cumodp_fftmul_uni(598, H, 299, F, 299, F, 1000) == cm.CUMODP_ASSUMPTION_ERROR
    # 257 = 2^8 + 1 cannot carry a 2^10-point transform
    assert cm.cumodp_fftmul_uni(598, H, 299, F, 299, F, 257) == \
        cm.CUMODP_FOURIER_DEGREE_TOO_SMALL
    # 2^27 points exceed the supported 2^26
    big = (1 << 26) + 1
    assert cm.lib().cumodp_fftmul_uni(big, ctypes.c_void_p(0), big // 2, ctypes.c_void_p(0),
                                      big // 2, ctypes.c_void_p(0), cm.P2) == \
        cm.CUMODP_FFT_SIZE_TOO_LARGE
    assert (H == -1).all()


def test_noalloc_multiply_rejects_before_device(cm):
    """The non-allocating multiply checks its limit (transform size 2^22),
    degrees and prime before any pointer is touched."""
    f = cm.lib().cumodp_fftmul_uni_dev_noalloc
    null = ctypes.c_void_p(0)
    half = 1 << 21
    assert f(2 * half, null, half, null, half, null, cm.P2, None) == cm.CUMODP_FFT_SIZE_TOO_LARGE
    assert b"2^22" in cm.lib().cumodp_last_error()
    assert f(10, null, -1, null, 5, null, cm.P2, None) == cm.CUMODP_ASSUMPTION_ERROR
    assert f(10, null, 5, null, 5, null, 1000, None) == cm.CUMODP_ASSUMPTION_ERROR


def test_oracle_is_not_linked_into_the_product(cm):
    r = subprocess.run(["ldd", cm.LIB_PATH], capture_output=True, text=True)
    assert "oracle" not in r.stdout
    raw, _ = exported(cm.LIB_PATH)
    assert not any(s.startswith("orc_") for s in raw)


# ---- host-side helpers of the integer multiplier (no device call) -------------
def _kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "tc_kats.json")) as f:
        return json.load(f)


def _words(vals):
    import oracle
    m = max(abs(int(v)) for v in vals)
    W = max(1, (m.bit_length() + 1 + 63) // 64)
    return oracle.ints_to_words([int(v) for v in vals], W)


def test_find_largest_bit_width_kats(cm):
    """unit_tests/test_two_convolution_poly_mul.cpp:9-54 through the library."""
    for c in _kats()["bit_width"]:
        a = _words([int(x) for x in c["a"]])
        b = _words([int(x) for x in c["b"]])
        got = cm.lib().cumodp_find_largest_bit_width(cm._ptr(a), a.shape[0], a.shape[1],
                                                     cm._ptr(b), b.shape[0], b.shape[1])
        assert got == c["bits"], c


def test_convert_to_modular_bivariate_kats(cm):
    """unit_tests/test_two_convolution_poly_mul.cpp:56-78 through the library."""
    for c in _kats()["limb_split"]:
        a = _words(c["a"])
        out = np.zeros(a.shape[0] * c["K"], np.int32)
        rc = cm.lib().cumodp_convert_to_modular_bivariate(cm._ptr(a), a.shape[0], a.shape[1],
                                                          c["K"], c["M"], c["prime"],
                                                          cm._ptr(out))
        assert rc == 0 and [int(x) for x in out] == c["expected"], c


def test_two_convolution_plan_choices(cm):
    # 34-bit coefficients of degree 16383: two primes, K = 2 limbs of 17 bits
    assert cm.two_convolution_plan(34, 16384, 16384) == (2, 2, 17)
    # 64-bit coefficients, 65536 terms: two primes, four 16-bit limbs
    assert cm.two_convolution_plan(64, 65536, 65536) == (2, 4, 16)
    # 1024-bit coefficients, 1024 terms
    npr, K, M = cm.two_convolution_plan(1024, 1024, 1024)
    assert K * M >= 1024 and M <= 30


def test_two_convolution_rejects_narrow_output(cm):
    a = np.array([2 ** 40, -3], np.int64)
    out = np.zeros(3, np.uint64)
    A = cm._words(a)
    rc = cm.lib().cumodp_two_convolution_poly_mul(cm._ptr(A), 2, 1, cm._ptr(A), 2, 1,
                                                  cm._ptr(out), 1)
    assert rc == cm.CUMODP_ASSUMPTION_ERROR


def test_gcd_and_division_link_stubs(cm):
    """gcdPrimeField / divPrimeField (include/cumodp.h:105-106) resolve for
    modpn's solver and report "not provided" without touching any output."""
    L = cm.lib()
    L.divPrimeField.restype = ctypes.c_float
    L.gcdPrimeField.restype = ctypes.c_float
    A = np.arange(1, 9, dtype=np.int32)
    B = np.array([3, 1], np.int32)
    R = np.full(1, -7, np.int32)
    Q = np.full(7, -7, np.int32)
    assert L.divPrimeField(cm._ptr(A), cm._ptr(B), cm._ptr(R), cm._ptr(Q), 8, 2, 469762049) == -1.0
    assert (R == -7).all() and (Q == -7).all()
    G = np.full(2, -7, np.int32)
    d = ctypes.c_int(-7)
    assert L.gcdPrimeField(ctypes.byref(d), cm._ptr(A), cm._ptr(B), cm._ptr(G), 8, 2, 469762049,
                           1) == -1.0
    assert d.value == -7 and (G == -7).all()
