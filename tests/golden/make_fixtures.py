"""Extract the reference's own known-answer vectors into committed fixtures.

Run in the development container only (it reads /root/reference, which does
not exist on GPU machines); the outputs are data files committed beside this
script:

* ``tc_kats.json`` -- the gtest known-answer tests of
  ``unit_tests/test_two_convolution_poly_mul.cpp`` (bit width :9-54, limb
  split :56-78, theta scaling :80-109, signed end-to-end products :112-192).
* ``bench_fixtures.npz`` -- the Maple-generated ``a, b, c = a*b`` triples of
  ``benchmark/*.test`` (three whitespace-separated integer lines each).  Each
  coefficient array is stored as (n, W) little-endian two's-complement uint64
  words.  For the four largest files only ``a``, ``b``, a SHA-256 of the
  canonical ``c`` words and 64 sampled coefficients of ``c`` are kept, to
  keep the repository small.

Usage:  python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import re

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

BIG_FILES = {"n1024.test", "n16384_m16384_p10000000000.test",
             "n8192_m8192_p10000000000.test", "n512.test"}


def words_needed(vals):
    m = max((abs(v) for v in vals), default=0)
    return max(1, (m.bit_length() + 1 + 63) // 64)


def to_words(vals, W):
    out = np.zeros((len(vals), W), np.uint64)
    mask = (1 << 64) - 1
    for i, v in enumerate(vals):
        v &= (1 << (64 * W)) - 1
        for j in range(W):
            out[i, j] = (v >> (64 * j)) & mask
    return out


def c_digest(cw):
    return hashlib.sha256(np.ascontiguousarray(cw, np.uint64).tobytes()).hexdigest()


def parse_int_list(s):
    out = []
    for m in re.finditer(r'mpz_class\("([0-9a-fA-F]+)",\s*(\d+)\)|(-?(?:0x[0-9a-fA-F]+|\d+))', s):
        if m.group(1):
            out.append(int(m.group(1), int(m.group(2))))
        else:
            out.append(int(m.group(3), 0))
    return out


def strip_comments(text):
    return "\n".join(l for l in text.splitlines() if not l.strip().startswith("//"))


def test_blocks(text):
    parts = re.split(r"TEST\(TwoConvolutionPolyMul,\s*(\w+)\)", text)
    return {parts[i]: strip_comments(parts[i + 1]) for i in range(1, len(parts), 2)}


def braces(s):
    """All top-level {...} groups in s, in order."""
    out, depth, start = [], 0, None
    for i, ch in enumerate(s):
        if ch == "{":
            if depth == 0:
                start = i
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                out.append(s[start + 1:i])
    return out


def kats():
    src = open(os.path.join(REF, "unit_tests/test_two_convolution_poly_mul.cpp")).read()
    blocks = test_blocks(src)
    res = {}
    # bit width: { a = {..}; b = {..}; run_test(a, b, N); }
    bw = []
    for sub in braces(braces(blocks["FindLargestBitWidthOfCoefficients"])[0]):
        if "run_test(a" not in sub:
            continue
        lists = re.findall(r"=\s*\{([^}]*)\}", sub)
        n = int(re.search(r"run_test\(a,\s*b,\s*(\d+)\)", sub).group(1))
        bw.append({"a": parse_int_list(lists[0]), "b": parse_int_list(lists[1]),
                   "bits": n})
    res["bit_width"] = [{"a": [str(v) for v in c["a"]], "b": [str(v) for v in c["b"]],
                         "bits": c["bits"]} for c in bw]
    # limb split
    ls = []
    for sub in braces(braces(blocks["ConvertToModularBivariate"])[0]):
        a = parse_int_list(re.search(r"a\s*=\s*\{([^}]*)\}", sub).group(1))
        N, K, M = map(int, re.search(r"\.N\s*=\s*(\d+),\s*\.K\s*=\s*(\d+),\s*\.M\s*=\s*(\d+)",
                                     sub).groups())
        prime = int(re.search(r"convert_to_modular_bivariate\(a,\s*base,\s*(\d+)\)", sub).group(1))
        exp = parse_int_list(re.search(r"BivariateMPZPolynomial\(\{([^}]*)\}\)", sub).group(1))
        ls.append({"a": a, "N": N, "K": K, "M": M, "prime": prime, "expected": exp})
    res["limb_split"] = ls
    # theta scaling
    sx = []
    for sub in braces(braces(blocks["ScaleXArgumentDev"])[0])[1:]:
        A = parse_int_list(re.search(r"A\s*=\s*\{([^}]*)\}", sub).group(1))
        N, K, M = map(int, re.search(r"\.N\s*=\s*(\d+),\s*\.K\s*=\s*(\d+),\s*\.M\s*=\s*(\d+)",
                                     sub).groups())
        theta = int(re.search(r"theta\s*=\s*(\d+)", sub).group(1))
        prime = int(re.search(r"prime\s*=\s*(\d+)", sub).group(1))
        exp = parse_int_list(re.search(r"\(BivariateMPZPolynomial\{([^}]*)\}\)", sub).group(1))
        sx.append({"A": A, "K": K, "theta": theta, "prime": prime, "expected": exp})
    res["scale_x"] = sx
    # end-to-end signed products
    e2e = []
    body = blocks["TwoConvolutionPolyMul"]
    for m in re.finditer(r"run_test\(\s*(\{[^;]*?\})\s*\);", body, re.S):
        groups = braces(m.group(1))
        if len(groups) != 3:
            continue
        a, b, c = (parse_int_list(g) for g in groups)
        e2e.append({"a": a, "b": b, "c": c})
    res["products"] = e2e
    return res


def bench_fixtures():
    d = os.path.join(REF, "benchmark")
    arrays = {}
    index = []
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".test"):
            continue
        lines = [l for l in open(os.path.join(d, fn)).read().splitlines() if l.strip()]
        a, b, c = ([int(t) for t in l.split()] for l in lines[:3])
        key = fn[:-5]
        wa, wb = words_needed(a), words_needed(b)
        wc = max(words_needed(c), wa + wb)
        arrays[key + "__a"] = to_words(a, wa)
        arrays[key + "__b"] = to_words(b, wb)
        cw = to_words(c, wc)
        entry = {"name": key, "na": len(a), "nb": len(b), "nc": len(c), "wa": wa, "wb": wb,
                 "wc": wc, "c_sha256": c_digest(cw), "max_bits": max(
                     max(abs(v) for v in a).bit_length(), max(abs(v) for v in b).bit_length())}
        if fn in BIG_FILES:
            rng = np.random.default_rng(12345)
            idx = np.sort(rng.choice(len(c), size=min(64, len(c)), replace=False))
            arrays[key + "__cidx"] = idx.astype(np.int64)
            arrays[key + "__csample"] = cw[idx]
            entry["full_c"] = False
        else:
            arrays[key + "__c"] = cw
            entry["full_c"] = True
        index.append(entry)
    return arrays, index


def main():
    k = kats()
    with open(os.path.join(HERE, "tc_kats.json"), "w") as f:
        json.dump(k, f, indent=1)
    arrays, index = bench_fixtures()
    np.savez_compressed(os.path.join(HERE, "bench_fixtures.npz"), **arrays)
    with open(os.path.join(HERE, "bench_fixtures.json"), "w") as f:
        json.dump(index, f, indent=1)
    print("kats:", {key: len(v) for key, v in k.items()})
    print("bench fixtures:", len(index))


if __name__ == "__main__":
    main()
