"""Exchange probe: LDS against cross-lane exchanges between the register
rounds of an NTT (kernels in bench_exchange.hip).

Three kernels run the same 1024-point forward transform per wavefront (16
words per lane, rounds of 4, 4 and 2 stages) and differ only in how the
elements change hands between rounds:
  lds  through the wavefront's LDS region,
  xt   by cross-lane transposes (v_permlane32/16_swap, DPP moves),
  xb   by cross-lane butterflies (each lane fetches its partner's word and
       computes both outputs of the pair).

    python bench_exchange.py --build     compile libbench_exchange.so (gfx950)
    python bench_exchange.py             check and time the three kernels

The check (one transform per vector) compares every kernel's output with
the others over the whole buffer and with a direct DFT of a few vectors;
the timing runs `--reps` transforms per vector between one load and one
store, on the null stream, timed with events around `--iters` launches.
Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(ROOT, "bench_exchange.hip")
LIB = os.path.join(ROOT, "libbench_exchange.so")
P = 2013265921            # 15 * 2^27 + 1
VARIANTS = ("lds", "xt", "xb")
WORDS_PER_BLOCK = 4 * 1024


def build():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
           "-shared", "-Wno-unused-command-line-argument",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", ROOT, SRC, "-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def primitive_root(p):
    n, fs, f = p - 1, [], 2
    while f * f <= n:
        if n % f == 0:
            fs.append(f)
            while n % f == 0:
                n //= f
        f += 1
    if n > 1:
        fs.append(n)
    g = 2
    while any(pow(g, (p - 1) // q, p) == 1 for q in fs):
        g += 1
    return g


def level_table(p, logn):
    """Levels M = 2 .. 2^logn: w_M^j 2^32 mod p at offset M/2 - 1 + j, with
    w_M = w^(N/M) for the N-point root w; returns (table, w)."""
    import numpy as np
    N = 1 << logn
    w = pow(primitive_root(p), (p - 1) // N, p)
    tab = np.zeros(N - 1, np.uint32)
    for lg in range(1, logn + 1):
        M = 1 << lg
        wm = pow(w, N // M, p)
        x = 1
        for j in range(M // 2):
            tab[M // 2 - 1 + j] = (x << 32) % p
            x = x * wm % p
    return tab, w


def dft_bitrev(x, w, p):
    """Y[brev(j)] = sum_i x_i w^(ij): the order a DIF transform leaves."""
    import numpy as np
    N = len(x)
    logn = N.bit_length() - 1
    pw = np.array([pow(w, t, p) for t in range(N)], dtype=np.int64)
    idx = np.outer(np.arange(N), np.arange(N)) % N
    Y = ((pw[idx] * x.astype(np.int64)[None, :]) % p).sum(axis=1) % p
    rev = np.array([int(format(j, f"0{logn}b")[::-1], 2) for j in range(N)])
    out = np.empty(N, np.int64)
    out[rev] = Y
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--build", action="store_true")
    ap.add_argument("--blocks", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=64)
    ap.add_argument("--iters", type=int, default=10)
    args = ap.parse_args()
    if args.build:
        print(build())
        return
    import numpy as np
    import torch
    if not os.path.exists(LIB):
        raise SystemExit(f"{LIB} is missing: run python bench_exchange.py --build")
    lib = ctypes.CDLL(LIB)
    lib.probe_launch.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                                 ctypes.c_uint, ctypes.c_int, ctypes.c_int]
    np_ = (-pow(P, -1, 1 << 32)) % (1 << 32)
    tab, w = level_table(P, 10)
    dev = torch.device("cuda:0")
    lv = torch.from_numpy(tab.view(np.int32)).to(dev)
    rng = np.random.default_rng(1)
    n = args.blocks * WORDS_PER_BLOCK
    x0 = rng.integers(0, P, n, dtype=np.int64).astype(np.uint32)
    X0 = torch.from_numpy(x0.view(np.int32)).to(dev)

    def launch(v, X, reps):
        rc = lib.probe_launch(v, X.data_ptr(), lv.data_ptr(), P, np_, reps, args.blocks)
        if rc != 0:
            raise RuntimeError(f"probe launch {VARIANTS[v]}: hip error {rc}")

    # check: one transform, all kernels agree, first vectors match a DFT
    outs = []
    for v in range(len(VARIANTS)):
        X = X0.clone()
        launch(v, X, 1)
        torch.cuda.synchronize()
        outs.append(X.cpu().numpy().view(np.uint32).astype(np.int64))
    for v in range(1, len(VARIANTS)):
        if not np.array_equal(outs[0], outs[v]):
            bad = int(np.nonzero(outs[0] != outs[v])[0][0])
            raise AssertionError(f"{VARIANTS[v]} differs from lds at word {bad}")
    for vec in range(4):
        sl = slice(vec * 1024, (vec + 1) * 1024)
        if not np.array_equal(outs[0][sl], dft_bitrev(x0[sl].astype(np.int64), w, P)):
            raise AssertionError(f"vector {vec} differs from the DFT")
    print(f"check: {len(VARIANTS)} kernels agree on {n} words; 4 vectors equal the DFT",
          file=sys.stderr, flush=True)

    res = {}
    X = X0.clone()
    for v, name in enumerate(VARIANTS):
        launch(v, X, args.reps)       # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.iters):
            launch(v, X, args.reps)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.iters
        transforms = args.blocks * 4 * args.reps
        res[name] = {"ms_per_launch": round(ms, 4),
                     "ns_per_1024pt_transform": round(ms * 1e6 / transforms, 4)}
        print(f"{name}: {ms:.4f} ms per launch", file=sys.stderr, flush=True)
    base = res["lds"]["ms_per_launch"]
    for name in VARIANTS:
        res[name]["vs_lds"] = round(res[name]["ms_per_launch"] / base, 3)
    print(json.dumps({"probe": "exchange", "p": P, "points": 1024, "words_per_lane": 16,
                      "blocks": args.blocks, "waves_per_block": 4, "reps": args.reps,
                      "iters": args.iters, "kernels": res}))


if __name__ == "__main__":
    main()
