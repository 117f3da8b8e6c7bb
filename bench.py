"""Benchmark of the NTT multiply path on one MI355X.

Default workload: one 2^24-point NTT multiply mod p = 2013265921 (two
degree 2^23-1 operands, uint32 residues uniform in [0, p) from SplitMix64
seed 2), inputs resident in HBM.  One step = one full product through
cumodp_fftmul_uni_dev (forward transforms of both operands, pointwise
product, inverse transform, scaling) on a single stream.

W warm-up steps are followed by untimed steps for RAMP_S seconds, which
bring the GPU from its idle clocks to its steady ones; then K steps are timed
twice: once as a caller runs them (`value`), once with HIP events around
every kernel launch on its stream (the per-kernel durations).

Prints one JSON line: metric/value (output coefficients per second over the
timed steps), the roofline of the dominant kernel (its algorithmic bytes per
launch over its average duration, measured with HIP events on its stream
during the second timed region), and the CPU baseline (the oracle's serial
NTT multiply on the same inputs, single thread, timed on this host).

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
# n = 2^24 points (4-byte residues): every kernel reads and writes each word
# it owns exactly once.
#   k_col_fwd: reads the lf + lg input words (the zero padding is not read),
#              writes 2n transformed words        -> 4 (lf + lg) + 8 n bytes
#   k_row_mul: F and G rows read (2n), product written (n) -> 12 n bytes
#   k_col_inv: reads n words, writes the nout product words -> 4 (n + nout)
def algo_bytes(name, n, lf, lg, nout):
    return {"k_col_fwd": 4 * (lf + lg) + 8 * n, "k_row_mul": 12 * n,
            "k_col_inv": 4 * (n + nout)}[name]


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


RAMP_S = 0.25
VALU_PEAK = 1024 * 2.32e9 / 4    # wave-instructions/s: 1024 SIMDs, 4 cycles each, 2.32 GHz
CPU_THREADS = 16                 # threads of the OpenMP CPU legs (the host-core share of one GPU)


def source_hash():
    """Hash of the library's kernel and host sources: counter profiles record
    the hash they were taken with, so counters from other sources show."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "cumodp_amd", "csrc")
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".h", ".cpp")):
            h.update(f.encode())
            with open(os.path.join(d, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def load_profile(name):
    """A committed counter summary (profiles/<name>) and whether it was
    taken from the current sources."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, False
    with open(path) as f:
        prof = json.load(f)
    return prof, prof.get("sources_sha16") == source_hash()


def valu_issue(line, seconds):
    """VALU issue fraction of one call of a secondary line: its SQ_INSTS_VALU
    per call (rocprofv3, profiles/valu_secondary.json) over the live time of
    the call, against the issue peak; None without a current profile."""
    prof, current = load_profile("valu_secondary.json")
    if not prof or line not in prof.get("lines", {}):
        return None
    insts = prof["lines"][line]["valu_insts_per_call"]
    out = {"insts_per_call": insts, "achieved": insts / seconds / 1e9,
           "peak": VALU_PEAK / 1e9, "unit": "G wave-instructions/s",
           "frac": insts / seconds / VALU_PEAK, "profile": "profiles/valu_secondary.json",
           "profile_from_current_sources": current}
    if not current:
        out["frac"] = None      # counters of other kernels: not reported as current
    return out


def ramp(fn, sync, seconds=RAMP_S):
    """Untimed repeats of fn() for at least `seconds` of wall time.  The GPU
    leaves its idle clocks only after tens of milliseconds of steady work:
    straight after an idle spell (a CPU leg, start-up) the first hundred
    2^24 multiplies run 10-15% slower than the steady state."""
    sync()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        sync()


def timed(fn, reps, sync):
    """Median wall time of fn() over reps runs, each bracketed by sync(),
    after the clock ramp."""
    ramp(fn, sync)
    ts = []
    for _ in range(reps):
        sync()
        t0 = time.perf_counter()
        fn()
        sync()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


GRAPHS = True   # --no-graph: the lines skip their graph capture and replays


def graph_replay(cm, fn_on, out, eager, reps):
    """Capture fn_on(stream) once on a created stream and time its replays
    (median of reps, after the clock ramp); returns (seconds, every replay
    equal to the eager output), or (None, None) under --no-graph."""
    import ctypes as C
    if not GRAPHS:
        return None, None
    L = cm.lib()
    st = L.cumodp_stream_create()
    try:
        g = cm.Graph.capture(st, lambda: fn_on(C.c_void_p(st)))
        sync = lambda: L.cumodp_stream_synchronize(C.c_void_p(st))
        t = timed(lambda: g.launch(st), reps, sync)
        sync()
        same = bool(np.array_equal(out.download(), eager))
        g.destroy()
    finally:
        L.cumodp_stream_destroy(C.c_void_p(st))
    return t, same


def bench_multiply_2048(cm, cpu, reps=200):
    """The 2048-point multiply mod 469762049 (two degree-1023 operands,
    SplitMix64 seed 1): the latency of one call (device buffers; and the host
    entry point cumodp_fftmul_uni with its copies), the latency of one forward
    2^11-point transform (stockham_dev on a device buffer, synchronous, the
    operation the reference publishes at 0.339872 ms on a Tesla M2050,
    new_tests/BigPrimeFieldFFT_3/test/timingReports/FFT_timings_TeslaM2050_node10:45),
    and the throughput of 4096 such products in one list call.  CPU legs: the
    oracle's schoolbook and serial NTT multiply on the same pair (one thread),
    and the 4096 products on one thread and on CPU_THREADS OpenMP threads."""
    import ctypes as C
    w = cm.workload1()
    p, F, G = w["p"], w["F"], w["G"]
    L = cm.lib()
    nout = len(F) + len(G) - 1
    dF, dG, dH = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G), cm.DeviceArray(nout)
    st = L.cumodp_stream_create()
    sync = lambda: L.cumodp_stream_synchronize(C.c_void_p(st))
    one = lambda: cm.fftmul_uni_dev(nout - 1, dH, len(F) - 1, dF, len(G) - 1, dG, p, st)
    one()
    t_dev = timed(one, reps, sync)
    H = dH.download()
    Hh = np.zeros(nout, np.int32)
    host = lambda: cm._check("cumodp_fftmul_uni", cm.cumodp_fftmul_uni(nout - 1, Hh, len(F) - 1,
                                                                       F, len(G) - 1, G, p))
    host()
    t_host = timed(host, reps, lambda: None)
    # one forward 2^11 transform in place on a device buffer (synchronous)
    X = np.concatenate([F, np.zeros(1024, np.int32)])
    dX = cm.DeviceArray.from_host(X)
    w11 = next(x for x in (pow(g, (p - 1) >> 11, p) for g in range(2, 64))
               if pow(x, 1 << 10, p) != 1)          # a root of order exactly 2^11
    fwd = cm.cxx("stockham_dev_w")
    fwd.restype = None
    t_fwd = timed(lambda: fwd(C.c_void_p(dX.ptr), 2048, 11, w11, p), reps, lambda: None)
    dX.upload(X)               # the timed calls transformed it repeatedly in place
    fwd(C.c_void_p(dX.ptr), 2048, 11, w11, p)
    Y = dX.download()
    # 4096 independent degree-1023 pairs in one batched call
    pairs = 4096
    rng = cm.SplitMix64(101)
    FF = cm.uniform_mod(rng, pairs * 1024, p)
    GG = cm.uniform_mod(rng, pairs * 1024, p)
    items, nf, ng, nh = cm.make_items(np.full(pairs, 1023), np.full(pairs, 1023))
    dFF, dGG, dHH = cm.DeviceArray.from_host(FF), cm.DeviceArray.from_host(GG), cm.DeviceArray(nh)
    batch = lambda: cm.list_fftmul_dev(items, dFF, dGG, dHH, p, st)
    batch()
    t_batch = timed(batch, 20, sync)
    HH = dHH.download()
    L.cumodp_stream_destroy(C.c_void_p(st))
    for d in (dF, dG, dH, dX, dFF, dGG, dHH):
        d.free()
    out = {"workload": "two degree-1023 polynomials mod 469762049, NTT size 2048",
           "latency_us_device": t_dev * 1e6, "latency_us_host_api_with_copies": t_host * 1e6,
           "forward_ntt_2e11_us": t_fwd * 1e6,
           "published_forward_ntt_2e11_us_tesla_m2050": 339.872,
           "batched_4096_pairs": {"ms": t_batch * 1e3, "coeffs_per_s": nh / t_batch,
                                  "products_per_s": pairs / t_batch},
           "check_host_eq_device": bool(np.array_equal(H, Hh)),
           "valu_issue_batched": valu_issue("multiply_2048", t_batch)}
    if cpu:
        import oracle
        oracle.lib()
        out["check_bit_exact_vs_oracle"] = bool(np.array_equal(H, oracle.plain_polymul(F, G, p)))
        out["check_forward_ntt_vs_oracle"] = bool(np.array_equal(Y, oracle.stockham(X, w11, p)))
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.plain_polymul(F, G, p)
            ts.append(time.perf_counter() - t0)
        t_school = float(np.median(ts))
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.ntt_polymul(F, G, p)
            ts.append(time.perf_counter() - t0)
        t_ntt = float(np.median(ts))
        desc = np.array([[it.f_off, it.g_off, it.h_off, it.df, it.dg] for it in items], np.int64)
        t0 = time.perf_counter()
        H1 = oracle.polymul_items(desc, FF, GG, nh, p, plain_cut=0, threads=1)
        t_1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        Hn = oracle.polymul_items(desc, FF, GG, nh, p, plain_cut=0, threads=CPU_THREADS)
        t_n = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": nh / t_1, "unit": "coeffs/s", "cores": 1, "kind": "port",
            "sample": "the 4096 pairs, one serial oracle NTT multiply each",
            "schoolbook_us_per_product": t_school * 1e6, "ntt_us_per_product": t_ntt * 1e6,
            "openmp": {"threads": CPU_THREADS, "nproc": os.cpu_count(),
                       "affinity_cores": len(os.sched_getaffinity(0)),
                       "coeffs_per_s": nh / t_n},
            "check_bit_exact": bool(np.array_equal(H1, HH)) and bool(np.array_equal(Hn, HH))}
    return out


def bench_list_multiply(cm, cpu, reps=5):
    """4096 pairs of uneven degree (log-uniform in [32, 16384]) mod 469762049:
    one batched cumodp_list_fftmul_dev call against one cumodp_fftmul_uni_dev
    call per pair, on the same resident inputs.  CPU leg: the oracle's serial
    NTT multiply over a bounded prefix of the pairs, checked against the
    batched outputs."""
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
    # the same batched call captured once into a graph and replayed
    L = cm.lib()
    tg, H3 = None, None
    if GRAPHS:
        st = L.cumodp_stream_create()
        dH3 = cm.DeviceArray(nh)
        g = cm.Graph.capture(st, lambda: cm.list_fftmul_dev(items, dF, dG, dH3, p, st))
        tg = timed(lambda: g.launch(st), reps, lambda: L.cumodp_stream_synchronize(st))
        H3 = dH3.download()
        g.destroy()
        L.cumodp_stream_destroy(st)
        dH3.free()
    for d in (dF, dG, dH1, dH2):
        d.free()
    out = {"workload": "4096 pairs, degrees log-uniform [32, 16384], mod 469762049",
           "output_coeffs": nh,
           "batched": {"ms": tb * 1e3, "coeffs_per_s": nh / tb},
           "per_call": {"ms": tc * 1e3, "coeffs_per_s": nh / tc},
           "graph_replay": tg and {"ms": tg * 1e3, "coeffs_per_s": nh / tg},
           "batched_over_per_call": tc / tb,
           "valu_issue_batched": valu_issue("list_multiply", tb),
           "check_batched_eq_per_call": bool(np.array_equal(H1, H2)),
           "check_graph_eq_batched": None if H3 is None else bool(np.array_equal(H1, H3)),
           "tuning": {"list_persistent": L.cumodp_tune_get_list_persistent(),
                      "list_tile_words": L.cumodp_tune_get_list_tile(),
                      "list_plain_cutoff": L.cumodp_tune_get_list_plain()}}
    if cpu:
        # the reference's own CPU multiply (modpn EX_Mont_Mul, built from its
        # sources into oracle/_ref) where it was built, else the oracle
        from oracle import modpn_ref
        if modpn_ref.available():
            mul, kind, what = modpn_ref.mul, "reference", "modpn EX_Mont_Mul (plain below " \
                "degree 138, Montgomery FFT above; modpn/src/FMUL.c:2046-2070)"
        else:
            import oracle
            mul, kind, what = oracle.ntt_polymul, "port", "serial oracle NTT multiply"
        t0, k, coeffs, ok = time.perf_counter(), 0, 0, True
        while k < len(items) and time.perf_counter() - t0 < 8.0:
            it = items[k]
            f = w["F"][it.f_off:it.f_off + it.df + 1]
            g = w["G"][it.g_off:it.g_off + it.dg + 1]
            h = mul(f, g, p)
            ok = ok and bool(np.array_equal(H1[it.h_off:it.h_off + it.df + it.dg + 1], h))
            coeffs += it.df + it.dg + 1
            k += 1
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": coeffs / t, "unit": "coeffs/s", "cores": 1, "kind": kind,
                               "sample": f"the first {k} pairs, one {what} per pair",
                               "check_bit_exact": ok}
    return out


def bench_plain_crossover(cm, cpu, reps=5):
    """Schoolbook (cumodp_list_plainmul_dev) against NTT (cumodp_list_fftmul_dev)
    list products of equal length N, 4096 pairs per size, mod 469762049: where
    the plain multiply stops winning.  CPU leg: modpn's plain multiply (its
    EX_Mont_Mul below degree 138) where oracle/_ref was built, else the
    oracle's schoolbook, on a sample of the N = 64 pairs."""
    p = 469762049
    pairs = 4096
    rng = cm.SplitMix64(77)
    out = {"workload": "4096 pairs of equal length N mod 469762049, plain vs NTT list multiply",
           "sizes": {}}
    L = cm.lib()
    for N in (16, 32, 64, 128, 256, 512):
        F = cm.uniform_mod(rng, pairs * N, p)
        G = cm.uniform_mod(rng, pairs * N, p)
        items, nf, ng, nh = cm.make_items(np.full(pairs, N - 1), np.full(pairs, N - 1))
        dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
        dH1, dH2 = cm.DeviceArray(nh), cm.DeviceArray(nh)
        run_p = lambda: cm.list_plainmul_dev(items, dF, dG, dH1, p)
        run_n = lambda: cm.list_fftmul_dev(items, dF, dG, dH2, p)
        cut = L.cumodp_tune_get_list_plain()
        L.cumodp_tune_set_list_plain(0)          # the NTT leg: every product on the NTT
        try:
            run_p(), run_n()
            tp = timed(run_p, reps, cm.synchronize)
            tn = timed(run_n, reps, cm.synchronize)
        finally:
            L.cumodp_tune_set_list_plain(cut)
        H1, H2 = dH1.download(), dH2.download()
        for d in (dF, dG, dH1, dH2):
            d.free()
        out["sizes"][str(N)] = {"plain_ms": tp * 1e3, "ntt_ms": tn * 1e3,
                                "plain_over_ntt_speed": tn / tp,
                                "check_plain_eq_ntt": bool(np.array_equal(H1, H2))}
        if N == 64:
            keep = (F, G, H1, items)
    won = [int(n) for n, v in out["sizes"].items() if v["plain_over_ntt_speed"] > 1.0]
    out["plain_faster_up_to_N"] = max(won) if won else None
    if cpu:
        from oracle import modpn_ref
        F, G, H1, items = keep
        if modpn_ref.available():
            mul, kind, what = modpn_ref.mul, "reference", "modpn EX_Mont_Mul (plain below degree 138)"
        else:
            import oracle
            mul, kind, what = oracle.plain_polymul, "port", "the oracle's schoolbook"
        t0, k, coeffs, ok = time.perf_counter(), 0, 0, True
        while k < len(items) and time.perf_counter() - t0 < 4.0:
            it = items[k]
            h = mul(F[it.f_off:it.f_off + 64], G[it.g_off:it.g_off + 64], p)
            ok = ok and bool(np.array_equal(H1[it.h_off:it.h_off + 127], h))
            coeffs += 127
            k += 1
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": coeffs / t, "unit": "coeffs/s", "cores": 1, "kind": kind,
                               "sample": f"{k} of the N = 64 pairs, {what} each",
                               "check_bit_exact": ok}
    return out


def bench_tree(cm, cpu, reps=5):
    """Subproduct tree over 2^16 points mod 469762049, all 17 levels.  CPU
    leg: the oracle's levels 1..8 (schoolbook per node), compared with the
    GPU's."""
    w = cm.workload4_tree()
    p, pts = w["p"], w["points"]
    k = 16
    L = cm.lib()
    nw = L.cumodp_subproduct_tree_words(k)
    dP, dT = cm.DeviceArray.from_host(pts), cm.DeviceArray(nw)
    run_on = lambda st: cm._check("cumodp_subproduct_tree_dev", L.cumodp_subproduct_tree_dev(
        dP.addr(), k, p, dT.addr(), st))
    run = lambda: run_on(None)
    run()
    t = timed(run, reps, cm.synchronize)
    tree = dT.download()
    tg, same_g = graph_replay(cm, run_on, dT, tree, reps)
    root = tree[L.cumodp_subproduct_tree_offset(k, k):]
    r = 123456789
    ok = eval_mod(root, r, p) == _prod_mod([(r - int(x)) % p for x in pts], p)
    # the plain per-call form: the same tree from the same level 0, one
    # cumodp_fftmul_uni_dev call per node product (2^16 - 1 calls), level by
    # level on one stream; timed once after one untimed run
    offs = [L.cumodp_subproduct_tree_offset(k, i) for i in range(k + 1)]
    dT2 = cm.DeviceArray.from_host(np.concatenate([tree[:offs[1]],
                                                   np.zeros(nw - offs[1], tree.dtype)]))
    base = dT2.addr().value

    def per_call():
        for i in range(1, k + 1):
            h = 1 << (i - 1)
            src, dst = base + 4 * offs[i - 1], base + 4 * offs[i]
            for j in range(1 << (k - i)):
                f = src + 4 * 2 * j * (h + 1)
                cm.fftmul_uni_dev(2 * h, dst + 4 * j * (2 * h + 1), h, f, h, f + 4 * (h + 1), p)

    per_call()
    cm.synchronize()
    t0 = time.perf_counter()
    per_call()
    cm.synchronize()
    tc = time.perf_counter() - t0
    same_pc = bool(np.array_equal(dT2.download(), tree))
    dP.free(), dT.free(), dT2.free()
    coeffs = sum((1 << (k - i)) * ((1 << i) + 1) for i in range(1, k + 1))
    out = {"workload": "subproduct tree, 2^16 points uniform mod 469762049, levels 0..16",
           "ms": t * 1e3, "graph_replay_ms": tg and tg * 1e3, "tree_words": nw,
           "coeffs_per_s": coeffs / t,
           "per_call": {"ms": tc * 1e3, "coeffs_per_s": coeffs / tc,
                        "calls": (1 << k) - 1},
           "batched_over_per_call": tc / t,
           "valu_issue": valu_issue("subproduct_tree", t),
           "check_root_evaluation": bool(ok), "check_graph_eq_eager": same_g,
           "check_per_call_eq_batched": same_pc}
    if cpu:
        import oracle
        oracle.lib()
        t0 = time.perf_counter()
        same = True
        for lev in range(1, 9):
            want = oracle.subproduct_level(pts, lev, p).ravel()
            off = L.cumodp_subproduct_tree_offset(k, lev)
            same = same and bool(np.array_equal(tree[off:off + want.size], want))
        tl = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": tl * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "levels 1..8 of the same tree by the oracle's "
                                         "schoolbook products", "check_levels_1_8": same}
    return out


def bench_eval(cm, cpu, reps=5):
    """Fast multipoint evaluation of a degree-(2^16-1) polynomial at the 2^16
    tree points mod 469762049 (tree, subinverse tree, remainder tree and
    Horner at the leaves).  CPU leg: the oracle's Horner on a sample of the
    points, checked against the GPU values."""
    w = cm.workload4_tree()
    p, pts = w["p"], w["points"]
    k = 16
    F = cm.uniform_mod(cm.SplitMix64(616), 1 << k, p)
    L = cm.lib()
    dF, dP, dV = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(pts), cm.DeviceArray(1 << k)
    run_on = lambda st: cm._check("cumodp_fast_evaluation_dev", L.cumodp_fast_evaluation_dev(
        dF.addr(), dP.addr(), k, p, dV.addr(), st))
    run = lambda: run_on(None)
    run()
    t = timed(run, reps, cm.synchronize)
    vals = dV.download()
    tg, same_g = graph_replay(cm, run_on, dV, vals, reps)
    for d in (dF, dP, dV):
        d.free()
    out = {"workload": "evaluate a degree-(2^16-1) polynomial at 2^16 points mod 469762049",
           "ms": t * 1e3, "points_per_s": (1 << k) / t, "graph_replay_ms": tg and tg * 1e3,
           "valu_issue": valu_issue("fast_evaluation", t),
           "graph_replay_points_per_s": tg and (1 << k) / tg, "check_graph_eq_eager": same_g}
    if cpu:
        import oracle
        oracle.lib()
        idx = np.arange(0, 1 << k, 64)
        t0 = time.perf_counter()
        want = oracle.poly_eval(F, pts[idx], p)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(idx) / tc, "unit": "points/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(idx)} of the points by the oracle's Horner "
                                         "rule (2^16 terms each)",
                               "check_bit_exact": bool(np.array_equal(vals[idx], want))}
    return out


def _prod_mod(vals, p):
    acc = 1
    for v in vals:
        acc = acc * v % p
    return acc


def bench_two_convolution(cm, cpu, reps=5):
    """Exact product of two degree-16383 int64 polynomials with coefficients
    uniform in (-10^10, 10^10) (34 bits); output as 2-word integers.  CPU leg:
    the oracle's two-convolution restatement and its exact schoolbook."""
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
    dA.free(), dB.free(), dC.free()
    out = {"workload": "two degree-16383 int64 polynomials, coefficients uniform in "
                       "(-10^10, 10^10), exact product",
           "plan": cm.two_convolution_plan(bits, na, nb),
           "ms": t * 1e3, "coeffs_per_s": nc / t,
           "valu_issue": valu_issue("two_convolution", t)}
    if cpu:
        import oracle
        oracle.lib()
        t0 = time.perf_counter()
        want = oracle.two_convolution_zmul(A, B, wc)
        t1 = time.perf_counter()
        school = oracle.zmul_i64(a, b)[:, :wc]
        t2 = time.perf_counter()
        out["cpu_baseline"] = {"value": nc / (t1 - t0), "unit": "coeffs/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full product by the oracle's two-convolution "
                                         "restatement (3 primes, K=32)",
                               "schoolbook_s": t2 - t1,
                               "check_bit_exact": bool(np.array_equal(Cw, want)) and
                               bool(np.array_equal(Cw, school))}
    return out


def _words_mod(Wd, q):
    """Two's-complement multi-word integers (rows of uint64 words) mod q."""
    Wd = np.asarray(Wd, np.uint64)
    acc = np.zeros(Wd.shape[0], np.uint64)
    qq = np.uint64(q)
    for j in range(Wd.shape[1]):
        f = np.uint64(pow(2, 64 * j, q))
        lo = Wd[:, j] & np.uint64(0xFFFFFFFF)
        hi = Wd[:, j] >> np.uint64(32)
        part = (lo % qq + (hi % qq) * np.uint64(pow(2, 32, q)) % qq) % qq
        acc = (acc + part * f % qq) % qq
    neg = (Wd[:, -1] >> np.uint64(63)).astype(bool)
    top = np.uint64(pow(2, 64 * Wd.shape[1], q))
    acc[neg] = (acc[neg] + qq - top) % qq
    return acc


def _eval_check(a, b, cw, x, q):
    """c(x) == a(x) b(x) mod q for int64 a, b and word rows cw."""
    def ev(v):
        v = np.asarray(v, np.uint64) % np.uint64(q)
        n = len(v)
        pw = np.empty(n, np.uint64)
        acc = 1
        blk = min(n, 4096)
        small = np.array([pow(x, j, q) for j in range(blk)], np.uint64)
        step = pow(x, blk, q)
        for s0 in range(0, n, blk):
            m = min(blk, n - s0)
            pw[s0:s0 + m] = small[:m] * np.uint64(acc) % np.uint64(q)
            acc = acc * step % q
        return int((v * pw % np.uint64(q)).sum() % np.uint64(q))
    am = np.asarray(a, np.int64) % q
    bm = np.asarray(b, np.int64) % q
    return ev(_words_mod(cw, q)) == ev(am) * ev(bm) % q


def bench_serving(cm, cpu, nreq=20000, max_batch_ms=1.5, age_factor=0.5):
    """Exact integer products served as they arrive: 20000 requests, degrees
    log-uniform in [64, 65536], coefficients of 16/32/48/64 bits (uniform
    choice), signed uniform.  First all requests arrive at once (capacity);
    then three runs of Poisson arrivals at 80% of that capacity (exponential
    gaps from the workload stream and from two more seeds), with batches
    capped at max_batch_ms of estimated GPU work, or at age_factor times the
    oldest waiting request's age when that is longer.  Latency = time from
    arrival to the product being back in host memory."""
    from cumodp_amd import serving
    t0 = time.perf_counter()
    w = cm.workload5(requests=nreq)
    Q = serving.make_requests(w)
    log(f"bench: serving: {Q.n} requests generated in {time.perf_counter() - t0:.1f} s, "
        f"{Q.total_coeffs} output coefficients")
    srv = serving.ZmulServer(slots=2, slot_bytes=1 << 30, age_factor=age_factor)
    runs = []
    try:
        srv.run(Q, np.zeros(Q.n))   # warm-up pass (plans, twiddles, buffers)
        done0, st0 = srv.run(Q, np.zeros(Q.n))
        cap_s = float(np.nanmax(done0))
        cap_rps = Q.n / cap_s
        lam = 0.8 * cap_rps
        # three Poisson arrival streams at 80% of capacity (the workload's own
        # gaps, then two more seeds); batches capped at max_batch_ms of
        # estimated GPU work as well as by footprint
        srv.max_batch_s = max_batch_ms * 1e-3
        for seed, gaps in ((5, w["exp_gaps"]), (51, cm.exp_gaps(51, Q.n)),
                           (52, cm.exp_gaps(52, Q.n))):
            arr = np.cumsum(gaps[:Q.n]) / lam
            done1, st1 = srv.run(Q, arr)
            span = float(np.nanmax(done1) - arr[0])
            runs.append({"gap_seed": seed, "coeffs_per_s": Q.total_coeffs / span,
                         **serving.latency_summary(done1, arr), **st1})
    finally:
        srv.close()
    lat = {"p50_ms": [r["p50_ms"] for r in runs], "p99_ms": [r["p99_ms"] for r in runs],
           "p99_over_p50": [r["p99_ms"] / r["p50_ms"] for r in runs]}
    # size-independent check on a sample: evaluation at a random point mod
    # two 31-bit primes
    idx = np.linspace(0, Q.n - 1, 200).astype(int)
    ok = all(_eval_check(Q.a(r), Q.b(r), Q.result(r), 987654321, q)
             for r in idx for q in (2147483647, 2013265921))
    out = {"workload": f"{Q.n} exact integer products, degrees log-uniform [64, 65536], "
                       "16/32/48/64-bit signed coefficients, Poisson arrivals at 80% of the "
                       "measured capacity",
           "output_coeffs": Q.total_coeffs,
           "capacity": {"s": cap_s, "coeffs_per_s": Q.total_coeffs / cap_s,
                        "requests_per_s": cap_rps, "valu_issue": valu_issue("serving", cap_s),
                        **st0},
           "poisson": {"rate_requests_per_s": lam, "max_batch_ms": max_batch_ms,
                       "age_factor": age_factor, **lat,
                       "runs": runs},
           "check_evaluation_sample": bool(ok)}
    if cpu:
        import oracle
        oracle.lib()
        t0, k, coeffs, same, alone = time.perf_counter(), 0, 0, True, True
        order = np.argsort(Q.nc)[::max(1, Q.n // 64)]
        for r in order:
            if time.perf_counter() - t0 > 10.0:
                break
            A = Q.a(r).view(np.uint64).reshape(-1, 1)
            B = Q.b(r).view(np.uint64).reshape(-1, 1)
            want = oracle.two_convolution_zmul(A, B, int(Q.wc[r]))
            same = same and bool(np.array_equal(Q.result(r), want))
            coeffs += int(Q.nc[r])
            k += 1
        t = time.perf_counter() - t0
        for r in order[:8]:
            one = cm.two_convolution_poly_mul(Q.a(r), Q.b(r), int(Q.wc[r]))
            alone = alone and bool(np.array_equal(one, Q.result(r)))
        out["cpu_baseline"] = {"value": coeffs / t, "unit": "coeffs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} requests spread over the size range, the "
                                         "oracle's two-convolution restatement each",
                               "check_bit_exact": same, "check_alone_eq_in_mix": alone}
    Q.free()
    return out


def copy_calibration(nbytes=256 << 20):
    """A plain device-to-device copy kernel of known size: the FETCH_SIZE and
    WRITE_SIZE counters of its dispatches calibrate the counters of the
    multiply kernels (bytes counted per byte moved)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"bench: copied {nbytes} bytes three times")


def counter_calls(cm, line, ncalls):
    """One warm-up call and `ncalls` more of a line's timed form (the batched
    call; for serving the capacity run over every request), with nothing else
    on the GPU.  A rocprofv3 --pmc run over this command gives the line's
    counters per call: bench_counters.py sums them over the dispatches (the
    runtime's own copy kernels aside) and divides by the calls printed here."""
    import ctypes as C
    L = cm.lib()
    free = []
    if line == "multiply_2048":
        p, pairs = 469762049, 4096
        rng = cm.SplitMix64(101)
        FF, GG = cm.uniform_mod(rng, pairs * 1024, p), cm.uniform_mod(rng, pairs * 1024, p)
        items, nf, ng, nh = cm.make_items(np.full(pairs, 1023), np.full(pairs, 1023))
        dF, dG, dH = cm.DeviceArray.from_host(FF), cm.DeviceArray.from_host(GG), cm.DeviceArray(nh)
        free += [dF, dG, dH]
        run = lambda: cm.list_fftmul_dev(items, dF, dG, dH, p)
    elif line == "list_multiply":
        w = cm.workload4()
        items, nf, ng, nh = cm.make_items(w["df"], w["dg"])
        dF, dG, dH = (cm.DeviceArray.from_host(w["F"]), cm.DeviceArray.from_host(w["G"]),
                      cm.DeviceArray(nh))
        free += [dF, dG, dH]
        run = lambda: cm.list_fftmul_dev(items, dF, dG, dH, w["p"])
    elif line in ("subproduct_tree", "fast_evaluation"):
        w = cm.workload4_tree()
        p, pts, k = w["p"], w["points"], 16
        dP = cm.DeviceArray.from_host(pts)
        free.append(dP)
        if line == "subproduct_tree":
            dT = cm.DeviceArray(L.cumodp_subproduct_tree_words(k))
            free.append(dT)
            run = lambda: cm._check("cumodp_subproduct_tree_dev", L.cumodp_subproduct_tree_dev(
                dP.addr(), k, p, dT.addr(), None))
        else:
            F = cm.uniform_mod(cm.SplitMix64(616), 1 << k, p)
            dF, dV = cm.DeviceArray.from_host(F), cm.DeviceArray(1 << k)
            free += [dF, dV]
            run = lambda: cm._check("cumodp_fast_evaluation_dev", L.cumodp_fast_evaluation_dev(
                dF.addr(), dP.addr(), k, p, dV.addr(), None))
    elif line == "two_convolution":
        w = cm.workload3()
        A, B = cm._words(w["a"], 1), cm._words(w["b"], 1)
        na, nb = len(w["a"]), len(w["b"])
        dA, dB = cm.DeviceArray.from_host(A.ravel()), cm.DeviceArray.from_host(B.ravel())
        dC = cm.DeviceArray((na + nb - 1) * 2, np.uint64)
        free += [dA, dB, dC]
        bits = L.cumodp_find_largest_bit_width(cm._ptr(A), na, 1, cm._ptr(B), nb, 1)
        run = lambda: cm._check("cumodp_two_convolution_poly_mul_dev",
                                L.cumodp_two_convolution_poly_mul_dev(
                                    dA.addr(), na, 1, dB.addr(), nb, 1, dC.addr(), 2, bits, None))
    elif line == "serving":
        from cumodp_amd import serving
        Q = serving.make_requests(cm.workload5())
        srv = serving.ZmulServer(slots=2, slot_bytes=1 << 30)   # the capacity run: no time cap
        run = lambda: srv.run(Q, np.zeros(Q.n))
    else:
        raise SystemExit(f"no counter form for line {line}")
    for i in range(1 + ncalls):
        run()
        cm.synchronize()
        log(f"bench: counter call {i + 1} of {1 + ncalls} ({line}) done")
    if line == "serving":
        srv.close()
        Q.free()
    for d in free:
        d.free()
    return {"line": line, "calls": 1 + ncalls, "sources_sha16": source_hash()}


SECONDARY = {
    "multiply_2048": bench_multiply_2048,
    "list_multiply": bench_list_multiply,
    "plain_crossover": bench_plain_crossover,
    "subproduct_tree": bench_tree,
    "fast_evaluation": bench_eval,
    "two_convolution": bench_two_convolution,
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--primary-only", action="store_true",
                    help="skip the list-multiply, tree and integer-multiply lines")
    ap.add_argument("--autotune", action="store_true",
                    help="measure and apply the tunables (column tile, four-step split, "
                         "list tile) before the timed runs")
    ap.add_argument("--serving-requests", type=int, default=20000,
                    help="requests of the serving line (0 skips it)")
    ap.add_argument("--serving-max-batch-ms", type=float, default=1.5,
                    help="service-time cap of a serving batch under Poisson arrivals")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (reference for "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--only", choices=sorted(SECONDARY) + ["serving"],
                    help="run one secondary line alone and print it (for profiling)")
    ap.add_argument("--counter-calls", type=int, default=0,
                    help="with --only: run only 1 + N calls of the line's timed form and print "
                         "the call count (the command profiled for counters per call)")
    ap.add_argument("--no-graph", action="store_true",
                    help="with --only: skip the graph capture and replays of the line")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path is single-GPU); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.only and args.counter_calls > 0:
        import cumodp_amd as cm
        print(json.dumps(counter_calls(cm, args.only, args.counter_calls)), flush=True)
        return
    if args.only:
        import cumodp_amd as cm
        global GRAPHS
        GRAPHS = not args.no_graph
        cpu = not args.no_cpu_baseline
        if args.only == "serving":
            r = bench_serving(cm, cpu, args.serving_requests, args.serving_max_batch_ms)
        else:
            r = SECONDARY[args.only](cm, cpu)
        print(json.dumps({args.only: r}), flush=True)
        return

    import cumodp_amd as cm
    L = cm.lib()
    tuning = None
    if args.autotune:
        from cumodp_amd import autotune
        tuning = autotune.tune_all(log=log)
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

    # the per-kernel timing events are pooled: warm-up fills the pool so the
    # timed regions create none
    cm.profile_enable(True)
    for _ in range(args.warmup):
        step()
    cm.synchronize()
    cm.profile_enable(False)
    # then untimed steps until the GPU runs at its steady clocks
    ramp(step, cm.synchronize)
    log("bench: warmup done")

    # timed region 1: the K steps as a caller runs them (no timing events
    # between the kernels); `value` comes from this one
    cm.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cm.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = nout / (ms_per_step * 1e-3)
    # timed region 2: the same K steps with a HIP event pair around every
    # kernel launch, on the stream the kernels run on; the per-kernel
    # durations (and the roofline's achieved rate) come from this one.  The
    # events cost about 10 us per step, so its step time is reported apart.
    cm.profile_enable(True)
    cm.profile_reset()
    cm.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cm.synchronize()
    t3 = time.perf_counter()
    cm.profile_enable(False)
    prof = cm.profile_read()
    ms_per_step_events = (t3 - t2) * 1e3 / args.steps
    log(f"bench: {ms_per_step:.4f} ms/step, {value:.4e} coeffs/s "
        f"({ms_per_step_events:.4f} ms/step with kernel events); kernels {prof}")

    H = dH.download()
    ok_eval = all(eval_mod(H, r, p) == eval_mod(F, r, p) * eval_mod(G, r, p) % p
                  for r in (3, 987654321))

    # dominant kernel (largest total time) and its roofline
    hot = ("k_col_fwd", "k_row_mul", "k_col_inv")
    name, (tot, cnt) = max(((k, v) for k, v in prof.items() if k in hot),
                           key=lambda kv: kv[1][0])
    avg_s = tot / cnt * 1e-3
    algo = algo_bytes(name, n, len(F), len(G), nout)
    achieved = algo / avg_s / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (calibrated on a copy kernel), kept in profiles/
    # and the VALU instructions per launch (SQ_INSTS_VALU): the issue rate,
    # not HBM, bounds these kernels
    # (reported only when the profile was taken from the current kernel
    # sources; a stale one is named with its figures set aside)
    traffic, counted = None, {}
    valu_peak = VALU_PEAK
    tj, current = load_profile("traffic_2e24.json")
    counter_src = {"profile": "profiles/traffic_2e24.json", "from_current_sources": current,
                   "sources_sha16": source_hash()}
    if tj and current:
        counted = tj["kernels"]
        valu_peak = tj.get("valu_peak_per_s", valu_peak)
        traffic = counted.get(name, {}).get("traffic_bytes")
    elif tj:
        counter_src["stale_traffic_bytes"] = tj["kernels"].get(name, {}).get("traffic_bytes")
    kernels = {}
    for k, v in prof.items():
        a = v[0] / v[1] * 1e-3
        kernels[k] = {"avg_us": round(a * 1e6, 2), "launches": v[1]}
        vi = counted.get(k, {}).get("valu_insts")
        if vi:
            kernels[k]["valu_issue_frac"] = round(vi / a / valu_peak, 3)
    vi = counted.get(name, {}).get("valu_insts")
    valu = None if not vi else {
        "insts_per_launch": vi, "achieved": vi / avg_s / 1e9, "peak": valu_peak / 1e9,
        "unit": "G wave-instructions/s", "frac": vi / avg_s / valu_peak}
    whole = 48 * n / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "poly-mul coeffs/s (bit-exact), single 2^24-point NTT multiply",
        "value": value,
        "unit": "coeffs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_ramp_s": RAMP_S,
        "ms_per_step": ms_per_step,
        "ms_per_step_with_kernel_events": ms_per_step_events,
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
                     "traffic": traffic, "counters": counter_src, "algo_bytes_per_launch": algo,
                     "avg_launch_us": avg_s * 1e6, "valu_issue": valu},
        "whole_multiply_hbm_frac_48n": whole / HBM_PEAK_GBS,
        "kernels": kernels,
        "tuning": {"autotuned": bool(args.autotune), "split_k2": L.cumodp_tune_get_split(24),
                   "col_tile_words": L.cumodp_tune_get_col_tile()},
        "check": {"evaluation_identity": ok_eval},
    }
    # the reference's boundary takes host buffers: the same multiply through
    # cumodp_fftmul_uni on host arrays (pageable memory), copies included;
    # reported beside `value`, never as it
    Hh = cm.poly_mul(F, G, p)
    th = []
    for _ in range(5):
        t0 = time.perf_counter()
        Hh = cm.poly_mul(F, G, p)
        th.append(time.perf_counter() - t0)
    t_host = float(np.median(th))
    out["host_api_with_copies"] = {
        "ms": t_host * 1e3, "coeffs_per_s": nout / t_host,
        "call": "cumodp_fftmul_uni on pageable host arrays: both operands in, the product out "
                "(2 x 32 MiB up, 64 MiB down), median of 5",
        "check_equals_device_result": bool(np.array_equal(Hh, H))}
    log(f"bench: host entry point with copies {t_host * 1e3:.2f} ms")
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle serial NTT multiply, 1 thread)")
        Hc, tc = cpu_baseline(F, G, p)
        out["check"]["bit_exact_vs_cpu"] = bool(np.array_equal(Hc, H))
        out["cpu_baseline"] = {"value": nout / tc, "unit": "coeffs/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full workload: one 2^24-point multiply, "
                                         "median of 2 runs of the oracle's serial NTT multiply",
                               "seconds_per_multiply": tc}
        from oracle import modpn_ref
        if modpn_ref.available():
            # modpn's Montgomery arithmetic is exact below 2^30 only, so the
            # reference's own CPU FFT multiply is timed at the same 2^24 size
            # mod 469762049 = 7*2^26+1, on inputs from the same generator
            log("bench: cpu reference (modpn CPU FFT multiply, 2^24 mod 469762049)")
            q = 469762049
            Fq, Gq = F % q, G % q
            t0 = time.perf_counter()
            Hq = modpn_ref.fft_mul(Fq, Gq, q)
            tq = time.perf_counter() - t0
            Hd = cm.poly_mul(Fq, Gq, q)
            out["cpu_reference"] = {
                "value": nout / tq, "unit": "coeffs/s", "cores": 1, "kind": "reference",
                "sample": "one 2^24-point multiply mod 469762049 by modpn's "
                          "EX_Mont_FFTMul_OPT2_AS_GENE (modpn/src/FMUL.c:1757), built from the "
                          "reference's C sources; modpn cannot run mod 2013265921",
                "seconds_per_multiply": tq,
                "check_bit_exact_vs_gpu": bool(np.array_equal(Hq, Hd))}
    L.cumodp_stream_destroy(stream)
    dF.free(), dG.free(), dH.free()
    if not args.primary_only:
        sec = {}
        cpu = not args.no_cpu_baseline
        for key, fn in SECONDARY.items():
            log(f"bench: {key}")
            sec[key] = fn(cm, cpu)
            log(f"bench: {key}: {sec[key]}")
        if args.serving_requests > 0:
            log("bench: serving")
            sec["serving"] = bench_serving(cm, cpu, args.serving_requests,
                                           args.serving_max_batch_ms)
            log(f"bench: serving: {sec['serving']}")
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
