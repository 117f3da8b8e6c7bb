"""Seeded synthetic workloads for the multiply path (shared by tests and bench).

Generator: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), one stream per workload, seed = workload
number.  Output i of a stream seeded with s is mix(s + i * 0x9E3779B97F4A7C15),
so a stream is generated vectorised with numpy.

Sampling rules:
* uniform in [0, p): the high 32 bits of a draw, rejected when >= the largest
  multiple of p below 2^32, then reduced mod p;
* signed uniform in (-B, B): a 64-bit draw, rejected when >= the largest
  multiple of 2B-1 below 2^64, then mapped to [-(B-1), B-1];
* log-uniform integer in [a, b]: floor(a * (b/a)^u) with u the top 53 bits of
  a draw as a fraction, clamped to [a, b].

Workloads (numbered as the benchmark configurations):
  1  two degree-1023 operands mod 469762049 (7*2^26+1), NTT size 2048
  2  two degree-(2^23-1) operands mod 2013265921 (15*2^27+1), NTT size 2^24
  3  two degree-16383 int64 operands uniform in (-10^10, 10^10)
  4  4096 pairs, degrees log-uniform in [32, 16384], mod 469762049; and a
     subproduct tree over 2^16 points uniform in [0, p)
  5  20000 integer requests: degrees log-uniform in [64, 65536], bit size
     uniform in {16, 32, 48, 64}, signed uniform coefficients
"""
import numpy as np

__all__ = ["SplitMix64", "uniform_mod", "signed_uniform", "log_uniform_int",
           "P1", "P2", "workload1", "workload2", "workload3", "workload4",
           "workload4_tree", "workload5", "exp_gaps"]

_GOLD = np.uint64(0x9E3779B97F4A7C15)
P1 = 469762049       # 7 * 2^26 + 1
P2 = 2013265921      # 15 * 2^27 + 1


def _mix(z):
    z = z.copy()
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed)
        self.count = 0

    def next(self, n):
        i = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = self.state + i * _GOLD
        self.count += n
        return _mix(z)


def _take(rng, need, accept):
    """Draw until `need` draws pass `accept`; unused draws go back to the stream."""
    out = []
    while need > 0:
        draws = rng.next(need + need // 16 + 16)
        idx = np.nonzero(accept(draws))[0][:need]
        consumed = idx[-1] + 1 if len(idx) == need else len(draws)
        rng.count -= len(draws) - int(consumed)
        out.append(draws[idx])
        need -= len(idx)
    return np.concatenate(out) if out else np.zeros(0, np.uint64)


def uniform_mod(rng, n, p):
    """n values uniform in [0, p) as int32."""
    limit = np.uint64(((1 << 32) // p) * p)
    z = _take(rng, n, lambda d: (d >> np.uint64(32)) < limit) >> np.uint64(32)
    return (z % np.uint64(p)).astype(np.int32)


def signed_uniform(rng, n, bound):
    """n int64 values uniform in (-bound, bound), bound <= 2^63."""
    R = 2 * bound - 1
    limit = ((1 << 64) // R) * R
    if limit >= (1 << 64):
        z = _take(rng, n, lambda d: np.ones(len(d), bool))
    else:
        z = _take(rng, n, lambda d: d < np.uint64(limit))
    with np.errstate(over="ignore"):
        v = (z % np.uint64(R)) - np.uint64(bound - 1)
    return v.view(np.int64)


def log_uniform_int(rng, n, a, b):
    z = rng.next(n)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    d = np.floor(a * np.power(b / a, u)).astype(np.int64)
    return np.clip(d, a, b)


def workload1():
    rng = SplitMix64(1)
    F = uniform_mod(rng, 1024, P1)
    G = uniform_mod(rng, 1024, P1)
    return {"p": P1, "F": F, "G": G}


def workload2(logn=24):
    rng = SplitMix64(2)
    half = 1 << (logn - 1)
    F = uniform_mod(rng, half, P2)
    G = uniform_mod(rng, half, P2)
    return {"p": P2, "F": F, "G": G}


def workload3(n=16384):
    rng = SplitMix64(3)
    a = signed_uniform(rng, n, 10 ** 10)
    b = signed_uniform(rng, n, 10 ** 10)
    return {"a": a, "b": b}


def workload4(pairs=4096, lo=32, hi=16384):
    rng = SplitMix64(4)
    df = log_uniform_int(rng, pairs, lo, hi)
    dg = log_uniform_int(rng, pairs, lo, hi)
    tot_f, tot_g = int((df + 1).sum()), int((dg + 1).sum())
    F = uniform_mod(rng, tot_f, P1)
    G = uniform_mod(rng, tot_g, P1)
    return {"p": P1, "df": df, "dg": dg, "F": F, "G": G}


def workload4_tree(logn=16):
    rng = SplitMix64(40)
    return {"p": P1, "points": uniform_mod(rng, 1 << logn, P1)}


def exp_gaps(seed, n):
    """n exponential inter-arrival gaps of mean 1 from SplitMix64(seed)."""
    rng = SplitMix64(seed)
    return -np.log1p(-((rng.next(n) >> np.uint64(11)).astype(np.float64) / (1 << 53)))


def workload5(requests=20000, lo=64, hi=65536):
    rng = SplitMix64(5)
    deg_a = log_uniform_int(rng, requests, lo, hi)
    deg_b = log_uniform_int(rng, requests, lo, hi)
    bits = np.array([16, 32, 48, 64])[(rng.next(requests) >> np.uint64(62)).astype(np.int64)]
    gaps = -np.log1p(-((rng.next(requests) >> np.uint64(11)).astype(np.float64) / (1 << 53)))
    return {"deg_a": deg_a, "deg_b": deg_b, "bits": bits, "exp_gaps": gaps, "rng": rng}
