"""Measured choice of the library's run-time tunables on the current GPU.

There are four tunables:

  * The four-step split of a 2^k-point transform (k > 13): rows of 2^k2
    words, with k2 in {12, 13, 14}, and columns of 2^(k-k2) words.
    `cumodp_tune_set_split`.
  * The tile of the list multiply's mixed-size kernel: 2048, 4096 or 8192
    words per workgroup, which sets how many small products share a
    workgroup and how many waves run per CU. `cumodp_tune_set_list_tile`.
  * The column tile of the four-step multiply: 16384 words with 1024
    threads, or 8192 words with 512 threads. The smaller tile gives more,
    smaller workgroups per CU and less lockstep between them, but each tile
    row is a shorter segment in memory. `cumodp_tune_set_col_tile`.
  * The plain cut-off of the list multiply: products whose shorter operand
    has at most this many coefficients run on the schoolbook kernel, the
    rest on the NTT. `cumodp_tune_set_list_plain`.
  * The list multiply's launch form: one persistent launch for every
    product up to 2^15 points, or one launch per size class (the mixed-size
    tile above applies to the latter). `cumodp_tune_set_list_persistent`.

Each candidate is timed on the operation it affects. The fastest candidate
is applied to the library, and the choices can be saved to and loaded from
a JSON file. The reference has no tuner; its launch shapes are fixed
constants (src/stockham.cu:40-46, src/list_stockham.cu:289-293).
"""
import ctypes as C
import json
import time

import numpy as np

from . import CumodpError, DeviceArray, _check, lib, list_fftmul_dev, make_items, synchronize
from .workload import P2, SplitMix64, uniform_mod, workload4


def _median_time(fn, reps, min_s=2e-3):
    """Median over reps of the time per call of fn, each rep running enough
    back-to-back calls (stream ordered, one synchronisation) to last min_s,
    so that launch and synchronisation latency do not decide the choice."""
    fn()
    synchronize()
    t0 = time.perf_counter()
    fn()
    synchronize()
    inner = max(1, int(min_s / max(time.perf_counter() - t0, 1e-6)) + 1)
    ts = []
    for _ in range(reps):
        synchronize()
        t0 = time.perf_counter()
        for _ in range(inner):
            fn()
        synchronize()
        ts.append((time.perf_counter() - t0) / inner)
    return float(np.median(ts))


def split_candidates(k):
    return [k2 for k2 in (12, 13, 14) if 1 <= k - k2 <= 13]


def tune_split(ks=range(14, 25), reps=5, p=P2, log=None):
    """Time one 2^k multiply (operands of 2^(k-1) words) for each split;
    apply and return {k: best k2}."""
    L = lib()
    rng = SplitMix64(77)
    best = {}
    for k in ks:
        half = 1 << (k - 1)
        F = uniform_mod(rng, half, p)
        G = uniform_mod(rng, half, p)
        dF, dG, dH = DeviceArray.from_host(F), DeviceArray.from_host(G), DeviceArray(2 * half - 1)
        fn = lambda: _check("cumodp_fftmul_uni_dev", L.cumodp_fftmul_uni_dev(
            2 * half - 2, dH.addr(), half - 1, dF.addr(), half - 1, dG.addr(), p, None))
        times = {}
        for k2 in split_candidates(k):
            _check("cumodp_tune_set_split", L.cumodp_tune_set_split(k, k2))
            times[k2] = _median_time(fn, reps)
        b = min(times, key=times.get)
        _check("cumodp_tune_set_split", L.cumodp_tune_set_split(k, b))
        best[k] = b
        if log:
            log(f"autotune: 2^{k} multiply: " +
                ", ".join(f"k2={c}: {t * 1e3:.3f} ms" for c, t in times.items()) + f" -> {b}")
        for d in (dF, dG, dH):
            d.free()
    return best


def tune_list_tile(reps=5, log=None, pairs=4096):
    """Time the uneven list multiply (the 4096-pair workload) for each tile
    size; apply and return the best."""
    L = lib()
    w = workload4(pairs=pairs)
    items, nf, ng, nh = make_items(w["df"], w["dg"])
    dF, dG, dH = DeviceArray.from_host(w["F"]), DeviceArray.from_host(w["G"]), DeviceArray(nh)
    fn = lambda: list_fftmul_dev(items, dF, dG, dH, w["p"])
    times = {}
    for words in (2048, 4096, 8192):
        _check("cumodp_tune_set_list_tile", L.cumodp_tune_set_list_tile(words))
        times[words] = _median_time(fn, reps)
    b = min(times, key=times.get)
    _check("cumodp_tune_set_list_tile", L.cumodp_tune_set_list_tile(b))
    if log:
        log("autotune: list multiply: " +
            ", ".join(f"{c} words: {t * 1e3:.3f} ms" for c, t in times.items()) + f" -> {b}")
    for d in (dF, dG, dH):
        d.free()
    return b


def tune_list_persistent(reps=5, log=None, pairs=4096):
    """Time the uneven list multiply (the 4096-pair workload) in its
    persistent and its per-class launch form; apply and return the best."""
    L = lib()
    w = workload4(pairs=pairs)
    items, nf, ng, nh = make_items(w["df"], w["dg"])
    dF, dG, dH = DeviceArray.from_host(w["F"]), DeviceArray.from_host(w["G"]), DeviceArray(nh)
    fn = lambda: list_fftmul_dev(items, dF, dG, dH, w["p"])
    times = {}
    for on in (1, 0):
        _check("cumodp_tune_set_list_persistent", L.cumodp_tune_set_list_persistent(on))
        times[on] = _median_time(fn, reps)
    b = min(times, key=times.get)
    _check("cumodp_tune_set_list_persistent", L.cumodp_tune_set_list_persistent(b))
    if log:
        log("autotune: list multiply launch: " +
            ", ".join(f"{'persistent' if c else 'per class'}: {t * 1e3:.3f} ms"
                      for c, t in times.items()) + f" -> {'persistent' if b else 'per class'}")
    for d in (dF, dG, dH):
        d.free()
    return b


def tune_list_plain(reps=5, log=None, pairs=4096, cutoffs=(0, 32, 64, 128, 192, 256, 384)):
    """Time the uneven list multiply (the 4096-pair workload) for each plain
    cut-off; apply and return the best."""
    L = lib()
    w = workload4(pairs=pairs)
    items, nf, ng, nh = make_items(w["df"], w["dg"])
    dF, dG, dH = DeviceArray.from_host(w["F"]), DeviceArray.from_host(w["G"]), DeviceArray(nh)
    fn = lambda: list_fftmul_dev(items, dF, dG, dH, w["p"])
    times = {}
    for c in cutoffs:
        _check("cumodp_tune_set_list_plain", L.cumodp_tune_set_list_plain(c))
        times[c] = _median_time(fn, reps)
    b = min(times, key=times.get)
    _check("cumodp_tune_set_list_plain", L.cumodp_tune_set_list_plain(b))
    if log:
        log("autotune: list multiply plain cut-off: " +
            ", ".join(f"{c}: {t * 1e3:.3f} ms" for c, t in times.items()) + f" -> {b}")
    for d in (dF, dG, dH):
        d.free()
    return b


def tune_col_tile(k=24, reps=5, p=P2, log=None):
    """Time the 2^k multiply with each column tile size; apply and return the
    best."""
    L = lib()
    rng = SplitMix64(78)
    half = 1 << (k - 1)
    F, G = uniform_mod(rng, half, p), uniform_mod(rng, half, p)
    dF, dG, dH = DeviceArray.from_host(F), DeviceArray.from_host(G), DeviceArray(2 * half - 1)
    fn = lambda: _check("cumodp_fftmul_uni_dev", L.cumodp_fftmul_uni_dev(
        2 * half - 2, dH.addr(), half - 1, dF.addr(), half - 1, dG.addr(), p, None))
    times = {}
    for words in (8192, 16384):
        _check("cumodp_tune_set_col_tile", L.cumodp_tune_set_col_tile(words))
        times[words] = _median_time(fn, reps)
    b = min(times, key=times.get)
    _check("cumodp_tune_set_col_tile", L.cumodp_tune_set_col_tile(b))
    if log:
        log(f"autotune: 2^{k} multiply column tile: " +
            ", ".join(f"{c} words: {t * 1e3:.3f} ms" for c, t in times.items()) + f" -> {b}")
    for d in (dF, dG, dH):
        d.free()
    return b


def tune_all(log=None, ks=range(14, 25)):
    L = lib()
    col = tune_col_tile(log=log)
    split = tune_split(ks=ks, log=log)
    # the mixed-size tile matters to the per-class form only
    on = L.cumodp_tune_get_list_persistent()
    _check("cumodp_tune_set_list_persistent", L.cumodp_tune_set_list_persistent(0))
    tile = tune_list_tile(log=log)
    _check("cumodp_tune_set_list_persistent", L.cumodp_tune_set_list_persistent(on))
    persistent = tune_list_persistent(log=log)
    return {"col_tile_words": col, "split": split, "list_tile_words": tile,
            "list_persistent": persistent, "list_plain_cutoff": tune_list_plain(log=log)}


def current():
    L = lib()
    return {"split": {k: L.cumodp_tune_get_split(k) for k in range(14, 27)},
            "list_tile_words": L.cumodp_tune_get_list_tile(),
            "col_tile_words": L.cumodp_tune_get_col_tile(),
            "list_plain_cutoff": L.cumodp_tune_get_list_plain(),
            "list_persistent": L.cumodp_tune_get_list_persistent()}


def apply(choices):
    L = lib()
    for k, k2 in choices.get("split", {}).items():
        _check("cumodp_tune_set_split", L.cumodp_tune_set_split(int(k), int(k2)))
    if "list_tile_words" in choices:
        _check("cumodp_tune_set_list_tile",
               L.cumodp_tune_set_list_tile(int(choices["list_tile_words"])))
    if "col_tile_words" in choices:
        _check("cumodp_tune_set_col_tile",
               L.cumodp_tune_set_col_tile(int(choices["col_tile_words"])))
    if "list_plain_cutoff" in choices:
        _check("cumodp_tune_set_list_plain",
               L.cumodp_tune_set_list_plain(int(choices["list_plain_cutoff"])))
    if "list_persistent" in choices:
        _check("cumodp_tune_set_list_persistent",
               L.cumodp_tune_set_list_persistent(int(choices["list_persistent"])))


def save(choices, path):
    with open(path, "w") as f:
        json.dump({"split": {str(k): v for k, v in choices["split"].items()},
                   "list_tile_words": choices["list_tile_words"],
                   "col_tile_words": choices.get("col_tile_words", 16384),
                   "list_plain_cutoff": choices.get("list_plain_cutoff", 32),
                   "list_persistent": choices.get("list_persistent", 1)}, f, indent=1)


def load(path):
    with open(path) as f:
        c = json.load(f)
    apply(c)
    return c
