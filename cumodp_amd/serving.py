"""Continuous-batching service for exact integer polynomial products.

Requests are products of two int64 polynomials (one word per coefficient).
Each request arrives at a given time with its data already in host memory.
The service batches whatever has arrived and runs each batch as one
cumodp_two_convolution_list_dev call. Inside the library a batch is split
into classes (transform size and limb plan), and each class runs as one
sequence of batched launches. Results are copied back into the caller's
result buffer.

Design:
  * Two slots, each with its own HIP stream and device buffers. While one
    batch computes, the next one uploads and launches on the other slot.
    Both streams have the default priority.
  * FIFO admission. A batch is a contiguous range of request ids, so its
    inputs and outputs are contiguous in the packed pinned host buffers:
    one host-to-device copy in and one device-to-host copy out per batch.
  * Admission by HBM footprint. A batch takes requests while its input,
    output and estimated transform scratch bytes stay under the slot's cap.
  * Admission by service time. A batch also stops growing once its
    estimated GPU time reaches `max_batch_s`. The estimate is a work count
    per request (transform scratch words times log2 of the y-transform
    length) times seconds per unit, learnt from the batches already served:
    each retired batch's service time (from the later of its launch and the
    previous batch's completion to its own completion) over its work count,
    as an exponentially weighted average. Without this cap a burst makes
    one batch of hundreds of requests, every request in it waits for all
    of them, and the arrivals during that long batch make the next one as
    long.  With `age_factor` a, the cap grows to a times the time the oldest
    waiting request has waited when that is longer: a backlog left by a
    burst then drains in fewer, larger (more efficient) batches instead of
    at the rate of capped batches, which barely exceeds the arrival rate.
  * Twiddle tables stay resident in the library across batches.

The reference has no serving layer. This drives the integer multiply that
replaces two_convolution_poly_mul (src/two_convolution_poly_mul.cu:532-569).
"""
import collections
import ctypes as C
import math
import time

import numpy as np

from . import CumodpError, ZmulItem, _check, lib


def out_words(bits, na, nb):
    """Words per product coefficient for operands of `bits` bits."""
    lg = max(0, math.ceil(math.log2(min(na, nb)))) if min(na, nb) > 1 else 0
    return (2 * bits + lg + 1 + 63) // 64


_ITEM_DTYPE = np.dtype([("a_off", np.int64), ("b_off", np.int64), ("c_off", np.int64),
                        ("na", np.int64), ("nb", np.int64), ("wa", np.int32), ("wb", np.int32),
                        ("wc", np.int32), ("bits", np.int32)])
assert _ITEM_DTYPE.itemsize == C.sizeof(ZmulItem)


class PinnedWords:
    """A pinned host buffer of uint64 words (hipHostMalloc) with a numpy view."""

    def __init__(self, nwords):
        self.n = max(int(nwords), 1)
        self.ptr = lib().cumodp_host_alloc(self.n * 8)
        if not self.ptr:
            raise CumodpError("cumodp_host_alloc", 0x401)
        self.words = np.ctypeslib.as_array((C.c_uint64 * self.n).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.words = None
            lib().cumodp_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Requests:
    """A packed set of requests: for request r, IN holds a (na words) then b
    (nb words) at in_off[r]; its product goes to OUT at out_off[r] as
    na+nb-1 coefficients of wc[r] words."""

    def __init__(self, a_list, b_list, bits):
        n = len(a_list)
        self.n = n
        self.na = np.array([len(a) for a in a_list], np.int64)
        self.nb = np.array([len(b) for b in b_list], np.int64)
        self.bits = np.asarray(bits, np.int64)
        self.wc = np.array([out_words(int(self.bits[r]), int(self.na[r]), int(self.nb[r]))
                            for r in range(n)], np.int64)
        self.nc = self.na + self.nb - 1
        self.in_off = np.zeros(n + 1, np.int64)
        self.in_off[1:] = np.cumsum(self.na + self.nb)
        self.out_off = np.zeros(n + 1, np.int64)
        self.out_off[1:] = np.cumsum(self.nc * self.wc)
        # per-request item fields (cumodp_zmul_item layout); offsets are made
        # relative to a batch's buffers when it launches
        self.items = np.zeros(n, _ITEM_DTYPE)
        self.items["a_off"] = self.in_off[:-1]
        self.items["b_off"] = self.in_off[:-1] + self.na
        self.items["c_off"] = self.out_off[:-1]
        self.items["na"], self.items["nb"] = self.na, self.nb
        self.items["wa"] = self.items["wb"] = 1
        self.items["wc"], self.items["bits"] = self.wc, self.bits
        self.IN = PinnedWords(int(self.in_off[-1]))
        self.OUT = PinnedWords(int(self.out_off[-1]))
        for r in range(n):
            o = int(self.in_off[r])
            self.IN.words[o:o + self.na[r]] = np.asarray(a_list[r], np.int64).view(np.uint64)
            o += int(self.na[r])
            self.IN.words[o:o + self.nb[r]] = np.asarray(b_list[r], np.int64).view(np.uint64)
        self.total_coeffs = int(self.nc.sum())

    def a(self, r):
        o = int(self.in_off[r])
        return self.IN.words[o:o + self.na[r]].view(np.int64)

    def b(self, r):
        o = int(self.in_off[r]) + int(self.na[r])
        return self.IN.words[o:o + self.nb[r]].view(np.int64)

    def result(self, r):
        """(nc, wc) two's-complement words of product r."""
        o = int(self.out_off[r])
        return self.OUT.words[o:o + self.nc[r] * self.wc[r]].reshape(-1, int(self.wc[r]))

    def free(self):
        self.IN.free()
        self.OUT.free()


class _Slot:
    def __init__(self):
        L = lib()
        self.stream = L.cumodp_stream_create()
        if not self.stream:
            raise CumodpError("cumodp_stream_create", 0x999)
        self.din = self.dout = None
        self.cap_in = self.cap_out = 0
        self.range = None            # (i0, i1) in flight
        self.units = 0.0             # its work count
        self.t_launch = 0.0          # when it was launched (seconds into the run)

    def ensure(self, nin, nout):
        L = lib()
        if nin > self.cap_in:
            if self.din:
                L.cumodp_dev_free(C.c_void_p(self.din))
            self.cap_in = max(nin, 2 * self.cap_in)
            self.din = L.cumodp_dev_alloc(self.cap_in * 8)
            if not self.din:
                raise CumodpError("cumodp_dev_alloc", 0x401)
        if nout > self.cap_out:
            if self.dout:
                L.cumodp_dev_free(C.c_void_p(self.dout))
            self.cap_out = max(nout, 2 * self.cap_out)
            self.dout = L.cumodp_dev_alloc(self.cap_out * 8)
            if not self.dout:
                raise CumodpError("cumodp_dev_alloc", 0x401)

    def close(self):
        L = lib()
        L.cumodp_stream_destroy(C.c_void_p(self.stream))
        for d in (self.din, self.dout):
            if d:
                L.cumodp_dev_free(C.c_void_p(d))
        self.din = self.dout = None


class ZmulServer:
    """Continuous batching over `slots` streams; a batch holds at most
    `slot_bytes` of inputs, outputs and estimated transform scratch, and at
    most max(`max_batch_s`, `age_factor` x the oldest waiting request's age)
    seconds of estimated GPU work (max_batch_s None: no time cap)."""

    def __init__(self, slots=2, slot_bytes=1 << 30, max_batch_s=None, age_factor=0.0):
        self.slots = [_Slot() for _ in range(slots)]
        self.slot_bytes = int(slot_bytes)
        self.max_batch_s = max_batch_s
        self.age_factor = age_factor
        self._age = 0.0   # how long the oldest waiting request has waited
        self.sec_per_unit = None     # learnt from served batches
        self._plans = {}

    def close(self):
        for s in self.slots:
            s.close()
        self.slots = []

    # scratch estimate of one request: 2 x 2K x Ly words per prime
    def _scratch_bytes(self, bits, ny):
        ey = max(4, (int(ny) - 1).bit_length())
        key = (int(bits), ey)
        v = self._plans.get(key)
        if v is None:
            npr, K, M = C.c_int(), C.c_int(), C.c_int()
            nm = 1 << (ey - 1)
            if not lib().cumodp_two_convolution_plan(int(bits), nm, nm, C.byref(npr), C.byref(K),
                                                     C.byref(M)):
                raise CumodpError("cumodp_two_convolution_plan", 0x202)
            v = 4 * 2 * (2 * K.value) * (1 << ey) * npr.value
            self._plans[key] = v
        return v

    def _units(self, bits, ny):
        """Work count of one request: its transform scratch words times log2
        of its y-transform length (proportional to its butterflies)."""
        ey = max(4, (int(ny) - 1).bit_length())
        return self._scratch_bytes(bits, ny) * ey / 4.0

    def _costs(self, Q):
        """Prefix sums over the requests of Q of their footprint bytes
        (inputs, outputs, estimated transform scratch) and work units, kept
        for the last Q seen: admission is then two binary searches, with no
        per-request Python work while the GPU waits."""
        if getattr(self, "_cost_q", None) is not Q:
            # y-transform exponent per request: max(4, bit_length(nc - 1))
            ey = np.maximum(4, np.frexp((Q.nc - 1).astype(np.float64))[1]).astype(np.int64)
            scratch = np.zeros(len(Q.nc), np.float64)
            for b, e in set(zip(Q.bits.tolist(), ey.tolist())):
                sel = (Q.bits == b) & (ey == e)
                scratch[sel] = self._scratch_bytes(b, (1 << (e - 1)) + 1)
            nbytes = 8.0 * (Q.na + Q.nb + Q.nc * Q.wc) + scratch
            units = scratch * ey / 4.0
            cb = np.zeros(len(Q.nc) + 1)
            cu = np.zeros(len(Q.nc) + 1)
            np.cumsum(nbytes, out=cb[1:])
            np.cumsum(units, out=cu[1:])
            self._cost_q, self._cost = Q, (cb, cu)
        return self._cost

    def _fit(self, Q, i0, i1):
        """Largest j in (i0, i1] such that requests [i0, j) fit one slot: the
        footprint cap and, once the rate is known, the service-time cap (the
        first request is always admitted)."""
        cb, cu = self._costs(Q)
        j = min(i1, int(np.searchsorted(cb, cb[i0] + self.slot_bytes, side="right")) - 1)
        if self.max_batch_s is not None and self.sec_per_unit:
            cap_s = max(self.max_batch_s, self.age_factor * self._age)
            cap_units = cap_s / self.sec_per_unit
            j = min(j, int(np.searchsorted(cu, cu[i0] + cap_units, side="right")) - 1)
        return max(j, i0 + 1)

    def _batch_units(self, Q, i0, i1):
        cu = self._costs(Q)[1]
        return float(cu[i1] - cu[i0])

    def _learn(self, units, seconds):
        """Fold one retired batch's measured service time into the rate."""
        if units <= 0 or seconds <= 0:
            return
        r = seconds / units
        self.sec_per_unit = r if self.sec_per_unit is None else 0.8 * self.sec_per_unit + 0.2 * r

    def _launch(self, slot, Q, i0, i1):
        L = lib()
        in0, in1 = int(Q.in_off[i0]), int(Q.in_off[i1])
        out0, out1 = int(Q.out_off[i0]), int(Q.out_off[i1])
        slot.ensure(in1 - in0, out1 - out0)
        st = C.c_void_p(slot.stream)
        _check("cumodp_memcpy_h2d_async", L.cumodp_memcpy_h2d_async(
            C.c_void_p(slot.din), C.c_void_p(Q.IN.ptr + 8 * in0), 8 * (in1 - in0), st))
        n = i1 - i0
        items = Q.items[i0:i1].copy()
        items["a_off"] -= in0
        items["b_off"] -= in0
        items["c_off"] -= out0
        _check("cumodp_two_convolution_list_dev", L.cumodp_two_convolution_list_dev(
            items.ctypes.data_as(C.POINTER(ZmulItem)), n, C.c_void_p(slot.din),
            C.c_void_p(slot.din), C.c_void_p(slot.dout), st))
        _check("cumodp_memcpy_d2h_async", L.cumodp_memcpy_d2h_async(
            C.c_void_p(Q.OUT.ptr + 8 * out0), C.c_void_p(slot.dout), 8 * (out1 - out0), st))
        slot.range = (i0, i1)
        slot.units = self._batch_units(Q, i0, i1)

    def run(self, Q, arrivals):
        """Serve requests 0..Q.n-1 arriving at `arrivals` (seconds from the
        start, non-decreasing).  Returns (done, stats): done[r] is the time
        (seconds from the start) at which product r was back in host memory."""
        L = lib()
        N = Q.n
        arrivals = np.asarray(arrivals, np.float64)
        done = np.full(N, np.nan)
        inflight = collections.deque()
        free = list(self.slots)
        head = 0
        batches = []
        t0 = time.perf_counter()
        retired = 0
        last_done = 0.0
        while retired < N:
            now = time.perf_counter() - t0
            # retire finished batches in launch order
            while inflight:
                q = L.cumodp_stream_query(C.c_void_p(inflight[0].stream))
                if q < 0:
                    raise CumodpError("cumodp_stream_query", 0x999)
                if q == 0:
                    break
                s = inflight.popleft()
                t = time.perf_counter() - t0
                i0, i1 = s.range
                done[i0:i1] = t
                retired += i1 - i0
                # its GPU time began when it was launched or when the batch
                # ahead of it finished, whichever was later
                self._learn(s.units, t - max(s.t_launch, last_done))
                last_done = t
                s.range = None
                free.append(s)
            arrived = int(np.searchsorted(arrivals, now, side="right"))
            if head < arrived and free:
                self._age = now - arrivals[head]
                j = self._fit(Q, head, arrived)
                s = free.pop(0)
                s.t_launch = time.perf_counter() - t0
                self._launch(s, Q, head, j)
                batches.append(j - head)
                inflight.append(s)
                head = j
                continue
            if not inflight and head < N and head >= arrived:
                # idle until the next arrival
                wait = arrivals[head] - (time.perf_counter() - t0)
                if wait > 2e-4:
                    time.sleep(wait - 1e-4)
        stats = {"batches": len(batches), "mean_batch": float(np.mean(batches)),
                 "max_batch": int(np.max(batches))}
        return done, stats


def make_requests(w, count=None):
    """Coefficients for a serving workload (cumodp_amd.workload5()): request
    r has deg_a[r]+1 and deg_b[r]+1 coefficients, signed uniform in
    (-2^(bits-1), 2^(bits-1)), drawn from the workload's stream in order."""
    from .workload import signed_uniform
    n = len(w["deg_a"]) if count is None else count
    rng = w["rng"]
    A, B = [], []
    for r in range(n):
        bound = 1 << (int(w["bits"][r]) - 1)
        A.append(signed_uniform(rng, int(w["deg_a"][r]) + 1, bound))
        B.append(signed_uniform(rng, int(w["deg_b"][r]) + 1, bound))
    return Requests(A, B, w["bits"][:n])


def latency_summary(done, arrivals):
    lat = (np.asarray(done) - np.asarray(arrivals)) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
