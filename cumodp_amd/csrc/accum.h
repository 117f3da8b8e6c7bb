// accum.h -- sums of products mod p in 64-bit accumulators.
//
// A sum of products of canonical residues (p < 2^31) is kept in a u64 and
// folded, acc -> hi * c + lo with c = 2^32 mod p, which is congruent mod p and
// at most B1 = (2^32 - 1)(c + 1).  Four products fit between folds for every
// p < 2^31: with k = floor(2^32 / p) >= 2 and c = 2^32 - k p,
// B1 + 4 (p-1)^2 = 2^64 - k 2^32 p + 4 p^2 + O(p) < 2^64, since 4 p < k 2^32
// (k = 2: p < 2^31; k >= 3: p <= 2^32 / k < k 2^30).  A value is reduced to
// [0, p) once at the end: two folds, one Montgomery reduction, one Montgomery
// product by 2^64 mod p.  Any odd modulus works (no root of unity needed).
#pragma once

#include "modarith.h"

namespace cm {

struct PlainMod {
    u32 p, np, c32, r2;   // modulus, -p^-1 mod 2^32, 2^32 mod p, 2^64 mod p
};

__device__ __forceinline__ u64 plain_fold(u64 acc, u32 c32) {
    return (u64)(u32)(acc >> 32) * c32 + (u32)acc;
}

// acc (any 64-bit value) -> acc mod p
__device__ __forceinline__ u32 plain_reduce(u64 acc, const PlainMod &m) {
    acc = plain_fold(plain_fold(acc, m.c32), m.c32);    // < 2^62 + 2^33
    const u32 q = (u32)acc * m.np;
    const u32 r = (u32)((acc + (u64)q * m.p) >> 32);    // acc / 2^32 mod p, < 2^32
    return cm_mul_tw(r, m.r2, m.p, m.np);               // r 2^64 / 2^32 = acc mod p
}


// sum_(k < cnt) x[k DX] y[k DY] mod p for canonical residues (DX, DY = +-1),
// four products per fold.  The four pairs of a group are loaded before any
// is multiplied, so their LDS or cache latencies overlap.
template <int DX, int DY>
__device__ __forceinline__ u32 plain_dot(const u32 *x, const u32 *y, int cnt,
                                         const PlainMod &m) {
    u64 acc = 0;
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
        const u32 x0 = x[k * DX], x1 = x[(k + 1) * DX], x2 = x[(k + 2) * DX], x3 = x[(k + 3) * DX];
        const u32 y0 = y[k * DY], y1 = y[(k + 1) * DY], y2 = y[(k + 2) * DY], y3 = y[(k + 3) * DY];
        acc += (u64)x0 * y0;
        acc += (u64)x1 * y1;
        acc += (u64)x2 * y2;
        acc += (u64)x3 * y3;
        acc = plain_fold(acc, m.c32);
    }
    for (; k < cnt; ++k) acc += (u64)x[k * DX] * y[k * DY];
    return plain_reduce(acc, m);
}

// Part `part` of PARTS of the same sum: indices [cnt part / PARTS,
// cnt (part + 1) / PARTS).  The parts' residues add up (mod p) to the sum.
template <int DX, int DY, int PARTS>
__device__ __forceinline__ u32 plain_dot_part(const u32 *x, const u32 *y, int cnt, int part,
                                              const PlainMod &m) {
    const int k0 = cnt * part / PARTS, k1 = cnt * (part + 1) / PARTS;
    return plain_dot<DX, DY>(x + k0 * DX, y + k0 * DY, k1 - k0, m);
}

// host: the constants for modulus p (odd, 3 <= p < 2^31)
static inline PlainMod plain_mod(u32 p) {
    PlainMod m;
    m.p = p;
    m.np = cm_host_npinv(p);
    m.c32 = (u32)(((u64)1 << 32) % p);
    m.r2 = (u32)((unsigned __int128)((u64)m.c32 * m.c32) % p);
    return m;
}

}  // namespace cm
