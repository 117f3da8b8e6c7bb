"""Benchmark of the NTT multiply path on one MI355X.

Default workload: one 2^24-point NTT multiply mod p = 2013265921 (two
degree 2^23-1 operands, uint32 residues uniform in [0, p) from SplitMix64
seed 2), inputs resident in HBM.  One step = one full product through
cumodp_fftmul_uni_dev (forward transforms of both operands, pointwise
product, inverse transform, scaling) on a single stream.

Prints one JSON line: metric/value (output coefficients per second over the
timed steps), the roofline of the dominant kernel (its algorithmic bytes per
launch over its average duration, measured with HIP events on its stream
during the timed region), and the CPU baseline (the oracle's serial NTT
multiply on the same inputs, single thread, timed on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s

# Algorithmic bytes per launch of each kernel of the four-step multiply of
# n = 2^24 words (4-byte residues): every kernel reads and writes each word
# it owns exactly once.
#   k_col_fwd: F and G columns, read 2n + write 2n words  -> 16 n bytes
#   k_row_mul: F and G rows read (2n), product written (n) -> 12 n bytes
#   k_col_inv: read n + write n words                     ->  8 n bytes
ALGO_BYTES_PER_WORD = {"k_col_fwd": 16, "k_row_mul": 12, "k_col_inv": 8}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def eval_mod(c, r, p):
    c = np.asarray(c, np.int64).astype(np.uint64)
    n = len(c)
    B = 4096
    small = np.empty(B, np.uint64)
    x = 1
    for j in range(B):
        small[j] = x
        x = x * r % p
    nb = (n + B - 1) // B
    big = np.empty(nb, np.uint64)
    y, rb = 1, pow(r, B, p)
    for a in range(nb):
        big[a] = y
        y = y * rb % p
    pw = ((big[:, None] * small[None, :]) % np.uint64(p)).ravel()[:n]
    return int(((c * pw) % np.uint64(p)).sum() % np.uint64(p))


def cpu_baseline(F, G, p, reps=2):
    """The oracle's serial NTT multiply (src/fft_aux.cu:862-879 restated)."""
    import oracle
    oracle.lib()
    times, H = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        H = oracle.ntt_polymul(F, G, p)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return H, t


def timed(fn, reps, sync):
    """Median wall time of fn() over reps runs, each bracketed by sync()."""
    ts = []
    for _ in range(reps):
        sync()
        t0 = time.perf_counter()
        fn()
        sync()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def bench_list_multiply(cm, reps=5):
    """4096 pairs of uneven degree (log-uniform in [32, 16384]) mod 469762049:
    one batched cumodp_list_fftmul_dev call against one cumodp_fftmul_uni_dev
    call per pair, on the same resident inputs."""
    w = cm.workload4()
    p = w["p"]
    items, nf, ng, nh = cm.make_items(w["df"], w["dg"])
    dF, dG = cm.DeviceArray.from_host(w["F"]), cm.DeviceArray.from_host(w["G"])
    dH1, dH2 = cm.DeviceArray(nh), cm.DeviceArray(nh)
    run_b = lambda: cm.list_fftmul_dev(items, dF, dG, dH1, p)
    run_c = lambda: cm.list_fftmul_dev(items, dF, dG, dH2, p, per_call=True)
    run_b(), run_c()
    tb = timed(run_b, reps, cm.synchronize)
    tc = timed(run_c, reps, cm.synchronize)
    H1, H2 = dH1.download(), dH2.download()
    # spot check three pairs against the oracle's schoolbook product
    import oracle
    ok = bool(np.array_equal(H1, H2))
    for i in (0, len(items) // 2, len(items) - 1):
        it = items[i]
        f = w["F"][it.f_off:it.f_off + it.df + 1]
        g = w["G"][it.g_off:it.g_off + it.dg + 1]
        ok = ok and bool(np.array_equal(H1[it.h_off:it.h_off + it.df + it.dg + 1],
                                        oracle.plain_polymul(f, g, p)))
    for d in (dF, dG, dH1, dH2):
        d.free()
    return {"workload": "4096 pairs, degrees log-uniform [32, 16384], mod 469762049",
            "output_coeffs": nh,
            "batched": {"ms": tb * 1e3, "coeffs_per_s": nh / tb},
            "per_call": {"ms": tc * 1e3, "coeffs_per_s": nh / tc},
            "batched_over_per_call": tc / tb, "check_batched_eq_per_call_eq_oracle": ok}


def bench_tree(cm, reps=5):
    """Subproduct tree over 2^16 points mod 469762049, all 17 levels."""
    w = cm.workload4_tree()
    p, pts = w["p"], w["points"]
    k = 16
    L = cm.lib()
    nw = L.cumodp_subproduct_tree_words(k)
    dP, dT = cm.DeviceArray.from_host(pts), cm.DeviceArray(nw)
    run = lambda: cm._check("cumodp_subproduct_tree_dev", L.cumodp_subproduct_tree_dev(
        dP.addr(), k, p, dT.addr(), None))
    run()
    t = timed(run, reps, cm.synchronize)
    T = dT.download(count=(1 << k) + 1, offset=L.cumodp_subproduct_tree_offset(k, k))
    r = 123456789
    ok = eval_mod(T, r, p) == _prod_mod([(r - int(x)) % p for x in pts], p)
    dP.free(), dT.free()
    return {"workload": "subproduct tree, 2^16 points uniform mod 469762049, levels 0..16",
            "ms": t * 1e3, "tree_words": nw, "check_root_evaluation": bool(ok)}


def _prod_mod(vals, p):
    acc = 1
    for v in vals:
        acc = acc * v % p
    return acc


def bench_two_convolution(cm, reps=5):
    """Exact product of two degree-16383 int64 polynomials with coefficients
    uniform in (-10^10, 10^10) (34 bits); output as 2-word integers."""
    w = cm.workload3()
    a, b = w["a"], w["b"]
    A, B = cm._words(a, 1), cm._words(b, 1)
    na, nb = len(a), len(b)
    nc, wc = na + nb - 1, 2
    L = cm.lib()
    dA, dB = cm.DeviceArray.from_host(A.ravel()), cm.DeviceArray.from_host(B.ravel())
    dC = cm.DeviceArray(nc * wc, np.uint64)
    bits = L.cumodp_find_largest_bit_width(cm._ptr(A), na, 1, cm._ptr(B), nb, 1)
    run = lambda: cm._check("cumodp_two_convolution_poly_mul_dev",
                            L.cumodp_two_convolution_poly_mul_dev(
                                dA.addr(), na, 1, dB.addr(), nb, 1, dC.addr(), wc, bits, None))
    run()
    t = timed(run, reps, cm.synchronize)
    Cw = dC.download().reshape(nc, wc)
    import oracle
    ok = bool(np.array_equal(Cw, oracle.zmul_i64(a, b)[:, :wc]))
    dA.free(), dB.free(), dC.free()
    return {"workload": "two degree-16383 int64 polynomials, coefficients uniform in "
                        "(-10^10, 10^10), exact product", "plan": cm.two_convolution_plan(
                            bits, na, nb),
            "ms": t * 1e3, "coeffs_per_s": nc / t, "check_exact_vs_schoolbook": ok}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--primary-only", action="store_true",
                    help="skip the list-multiply, tree and integer-multiply lines")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path is single-GPU); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import cumodp_amd as cm
    L = cm.lib()
    w = cm.workload2()
    F, G, p = w["F"], w["G"], w["p"]
    n = 1 << 24
    df, dg = len(F) - 1, len(G) - 1
    nout = df + dg + 1
    log(f"bench: workload 2^24 multiply mod {p}, steps={args.steps} warmup={args.warmup}")
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    dH = cm.DeviceArray(nout)
    stream = L.cumodp_stream_create()

    def step():
        cm.fftmul_uni_dev(nout - 1, dH, df, dF, dg, dG, p, stream)

    for _ in range(args.warmup):
        step()
    cm.synchronize()
    log("bench: warmup done")

    cm.profile_reset()
    cm.profile_enable(True)
    cm.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cm.synchronize()
    t1 = time.perf_counter()
    cm.profile_enable(False)
    prof = cm.profile_read()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = nout / (ms_per_step * 1e-3)
    log(f"bench: {ms_per_step:.4f} ms/step, {value:.4e} coeffs/s; kernels {prof}")

    H = dH.download()
    ok_eval = all(eval_mod(H, r, p) == eval_mod(F, r, p) * eval_mod(G, r, p) % p
                  for r in (3, 987654321))

    # dominant kernel (largest total time) and its roofline
    name, (tot, cnt) = max(((k, v) for k, v in prof.items() if k in ALGO_BYTES_PER_WORD),
                           key=lambda kv: kv[1][0])
    avg_s = tot / cnt * 1e-3
    algo = ALGO_BYTES_PER_WORD[name] * n
    achieved = algo / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_2e24.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(name)
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1]}
               for k, v in prof.items()}
    whole = 48 * n / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "poly-mul coeffs/s (bit-exact), single 2^24-point NTT multiply",
        "value": value,
        "unit": "coeffs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64 seed 2, uniform residues)",
        "config": {"workload": "one NTT multiply of two degree-(2^23-1) polynomials mod "
                               "2013265921, transform size 2^24 (four-step 2^11 columns x 2^13 rows)",
                   "n": n, "p": p},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algo_bytes_per_launch": algo,
                     "avg_launch_us": avg_s * 1e6},
        "whole_multiply_hbm_frac_48n": whole / HBM_PEAK_GBS,
        "kernels": kernels,
        "check": {"evaluation_identity": ok_eval},
    }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle serial NTT multiply, 1 thread)")
        Hc, tc = cpu_baseline(F, G, p)
        out["check"]["bit_exact_vs_cpu"] = bool(np.array_equal(Hc, H))
        out["cpu_baseline"] = {"value": nout / tc, "unit": "coeffs/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full workload: one 2^24-point multiply, "
                                         "median of 2 runs of the oracle's serial NTT multiply",
                               "seconds_per_multiply": tc}
    L.cumodp_stream_destroy(stream)
    dF.free(), dG.free(), dH.free()
    if not args.primary_only:
        sec = {}
        for key, fn in (("list_multiply", bench_list_multiply), ("subproduct_tree", bench_tree),
                        ("two_convolution", bench_two_convolution)):
            log(f"bench: {key}")
            sec[key] = fn(cm)
            log(f"bench: {key}: {sec[key]}")
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
