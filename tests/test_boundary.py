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


def test_oracle_is_not_linked_into_the_product(cm):
    r = subprocess.run(["ldd", cm.LIB_PATH], capture_output=True, text=True)
    assert "oracle" not in r.stdout
    raw, _ = exported(cm.LIB_PATH)
    assert not any(s.startswith("orc_") for s in raw)
