"""Graph capture of repeated shapes: every replay must equal the eager result.

Each graph is replayed several times. Between replays the outputs are
cleared and other eager work runs on another stream, with synchronisations.
Every replay is compared with the eager result of the same calls.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _replays(cm, stream, g, out, eager, other, n=3):
    L = cm.lib()
    for k in range(n):
        out.upload(np.zeros(out.n, out.dtype))
        other()
        cm.synchronize()
        g.launch(stream)
        assert L.cumodp_stream_synchronize(C.c_void_p(stream)) == 0
        other()
        cm.synchronize()
        assert np.array_equal(out.download(), eager), f"replay {k}"


def test_graph_list_multiply_and_four_step(cm, orc):
    L = cm.lib()
    w = cm.workload4(pairs=300, lo=32, hi=40000)
    p = w["p"]
    items, nf, ng, nh = cm.make_items(w["df"], w["dg"])
    dF, dG, dH = cm.DeviceArray.from_host(w["F"]), cm.DeviceArray.from_host(w["G"]), \
        cm.DeviceArray(nh)
    # unrelated eager work for between the replays: one 2^16 multiply
    x = cm.workload2(logn=16)
    dX, dY = cm.DeviceArray.from_host(x["F"]), cm.DeviceArray.from_host(x["G"])
    dZ = cm.DeviceArray(len(x["F"]) * 2 - 1)
    other = lambda: cm.fftmul_uni_dev(dZ.n - 1, dZ, len(x["F"]) - 1, dX, len(x["G"]) - 1, dY,
                                      x["p"])
    s = L.cumodp_stream_create()
    try:
        run = lambda: cm.list_fftmul_dev(items, dF, dG, dH, p, s)
        g = cm.Graph.capture(s, run)
        cm.synchronize()
        eager = dH.download()
        for i in (0, 1, len(items) // 2, len(items) - 1):
            it = items[i]
            f = w["F"][it.f_off:it.f_off + it.df + 1]
            gg = w["G"][it.g_off:it.g_off + it.dg + 1]
            assert np.array_equal(eager[it.h_off:it.h_off + it.df + it.dg + 1],
                                  orc.ntt_polymul(f, gg, p))
        _replays(cm, s, g, dH, eager, other)
        g.destroy()
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))


def test_graph_subproduct_tree(cm):
    L = cm.lib()
    w = cm.workload4_tree(logn=12)
    k, p = 12, w["p"]
    dP = cm.DeviceArray.from_host(w["points"])
    dT = cm.DeviceArray(L.cumodp_subproduct_tree_words(k))
    s = L.cumodp_stream_create()
    try:
        run = lambda: cm._check("cumodp_subproduct_tree_dev", L.cumodp_subproduct_tree_dev(
            dP.addr(), k, p, dT.addr(), C.c_void_p(s)))
        g = cm.Graph.capture(s, run)
        L.cumodp_stream_synchronize(C.c_void_p(s))
        eager = dT.download()
        _replays(cm, s, g, dT, eager, lambda: None)
        g.destroy()
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))


def test_graph_fast_evaluation(cm, orc):
    """Fast multipoint evaluation (tree, subinverse tree, remainder tree,
    Horner) captured once and replayed: every replay equals the eager
    values, which match Horner's rule on a sample of the points."""
    L = cm.lib()
    w = cm.workload4_tree(logn=12)
    k, p = 12, w["p"]
    F = cm.uniform_mod(cm.SplitMix64(77), 1 << k, p)
    dF, dP, dV = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(w["points"]), \
        cm.DeviceArray(1 << k)
    s = L.cumodp_stream_create()
    try:
        run = lambda: cm._check("cumodp_fast_evaluation_dev", L.cumodp_fast_evaluation_dev(
            dF.addr(), dP.addr(), k, p, dV.addr(), C.c_void_p(s)))
        g = cm.Graph.capture(s, run)
        L.cumodp_stream_synchronize(C.c_void_p(s))
        eager = dV.download()
        idx = np.array([0, 1, 777, (1 << k) - 1])
        assert np.array_equal(eager[idx], orc.poly_eval(F, w["points"][idx], p))
        _replays(cm, s, g, dV, eager, lambda: None)
        g.destroy()
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))


def test_graph_keeps_its_scratch(cm, orc):
    """A graph holds the scratch it was captured with: an eager call on the
    same stream that needs more scratch, and then the destruction of that
    stream, must leave the replays correct (they would read freed memory
    if the graph did not hold its buffers)."""
    L = cm.lib()
    small, big = cm.workload2(logn=15), cm.workload2(logn=20)
    dA, dB = cm.DeviceArray.from_host(small["F"]), cm.DeviceArray.from_host(small["G"])
    dH = cm.DeviceArray(2 * len(small["F"]) - 1)
    dX, dY = cm.DeviceArray.from_host(big["F"]), cm.DeviceArray.from_host(big["G"])
    dZ = cm.DeviceArray(2 * len(big["F"]) - 1)
    want = orc.ntt_polymul(small["F"], small["G"], small["p"])
    s = L.cumodp_stream_create()
    s2 = L.cumodp_stream_create()
    g = None
    try:
        run = lambda: cm.fftmul_uni_dev(dH.n - 1, dH, len(small["F"]) - 1, dA,
                                        len(small["G"]) - 1, dB, small["p"], s)
        g = cm.Graph.capture(s, run)
        # grows the stream's scratch past the captured size
        cm.fftmul_uni_dev(dZ.n - 1, dZ, len(big["F"]) - 1, dX, len(big["G"]) - 1, dY,
                          big["p"], s)
        assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
        dH.upload(np.zeros(dH.n, np.int32))
        g.launch(s)
        assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
        assert np.array_equal(dH.download(), want)
        assert L.cumodp_stream_destroy(C.c_void_p(s)) == 0
        s = None
        # the capture stream is gone; the graph replays on another stream
        for _ in range(2):
            dH.upload(np.zeros(dH.n, np.int32))
            g.launch(s2)
            assert L.cumodp_stream_synchronize(C.c_void_p(s2)) == 0
            assert np.array_equal(dH.download(), want)
        assert np.array_equal(dZ.download(), orc.ntt_polymul(big["F"], big["G"], big["p"]))
    finally:
        if g is not None:
            g.destroy()
        if s is not None:
            L.cumodp_stream_destroy(C.c_void_p(s))
        L.cumodp_stream_destroy(C.c_void_p(s2))


def test_graph_requires_created_stream(cm):
    assert cm.lib().cumodp_graph_begin(None) == cm.CUMODP_ASSUMPTION_ERROR


def test_autotune_choices_keep_results(cm, orc):
    """Every tunable setting gives the same products (tuning changes speed only)."""
    from cumodp_amd import autotune
    L = cm.lib()
    before = autotune.current()
    w = cm.workload4(pairs=200, lo=16, hi=9000)
    items, nf, ng, nh = cm.make_items(w["df"], w["dg"])
    dF, dG = cm.DeviceArray.from_host(w["F"]), cm.DeviceArray.from_host(w["G"])
    outs = []
    try:
        for words in (2048, 4096, 8192):
            assert L.cumodp_tune_set_list_tile(words) == 0
            dH = cm.DeviceArray(nh)
            cm.list_fftmul_dev(items, dF, dG, dH, w["p"])
            cm.synchronize()
            outs.append(dH.download())
        assert all(np.array_equal(outs[0], o) for o in outs[1:])
        it = items[len(items) - 1]
        f = w["F"][it.f_off:it.f_off + it.df + 1]
        g = w["G"][it.g_off:it.g_off + it.dg + 1]
        assert np.array_equal(outs[0][it.h_off:it.h_off + it.df + it.dg + 1],
                              orc.ntt_polymul(f, g, w["p"]))
        assert L.cumodp_tune_set_list_tile(1000) == cm.CUMODP_ASSUMPTION_ERROR
        # every plain cut-off (schoolbook for short operands) gives the same products
        for cut in (0, 16, 300, 1024):
            assert L.cumodp_tune_set_list_plain(cut) == 0
            dH = cm.DeviceArray(nh)
            cm.list_fftmul_dev(items, dF, dG, dH, w["p"])
            cm.synchronize()
            assert np.array_equal(outs[0], dH.download()), cut
        assert L.cumodp_tune_set_list_plain(2000) == cm.CUMODP_ASSUMPTION_ERROR
        # both column tiles give the same 2^20 product
        x = cm.workload2(logn=20)
        dX, dY = cm.DeviceArray.from_host(x["F"]), cm.DeviceArray.from_host(x["G"])
        prods = []
        for words in (16384, 8192):
            assert L.cumodp_tune_set_col_tile(words) == 0
            dZ = cm.DeviceArray(2 * len(x["F"]) - 1)
            cm.fftmul_uni_dev(dZ.n - 1, dZ, len(x["F"]) - 1, dX, len(x["G"]) - 1, dY, x["p"])
            cm.synchronize()
            prods.append(dZ.download())
        assert np.array_equal(prods[0], prods[1])
        assert np.array_equal(prods[0], orc.ntt_polymul(x["F"], x["G"], x["p"]))
        assert L.cumodp_tune_set_col_tile(4096) == cm.CUMODP_ASSUMPTION_ERROR
        best = autotune.tune_split(ks=[15], reps=2)
        assert best[15] in autotune.split_candidates(15)
        assert L.cumodp_tune_get_split(15) == best[15]
    finally:
        autotune.apply(before)
