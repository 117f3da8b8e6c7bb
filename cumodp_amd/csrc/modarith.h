// modarith.h -- exact arithmetic modulo a 32-bit Fourier prime p < 2^31.
//
// All residues are canonical ([0, p)) at function boundaries.  Three
// multiplication forms are used:
//   * Montgomery multiplication with R = 2^32 (P. Montgomery, 1985) in its
//     lazy form: for a * b < p * 2^32, t = a*b, m = t * (-p^-1) mod 2^32,
//     (t + m*p) / 2^32 lies in [0, 2p) and equals a*b/R mod p.  On gfx950
//     this is three full-rate instructions (v_mad_u64_u32 a*b, v_mul_lo_u32,
//     v_mad_u64_u32 m*p + t).  Transform twiddles and scale constants are
//     stored in Montgomery form (w R mod p), so one lazy product yields
//     a * w mod p in [0, 2p) for any 32-bit a.
//   * Shoup multiplication by a precomputed constant w with companion
//     wq = floor(w * 2^32 / p): q = mulhi(a, wq), r = a*w - q*p lies in
//     [0, 2p) for any 32-bit a (V. Shoup, NTL; D. Harvey, "Faster arithmetic
//     for number-theoretic transforms", 2014); used off the hot path.
//   * plain 64-bit remainder on the host for table construction.
// These replace the double-precision quotient estimate of the reference's
// mul_mod (include/inlines.h:95-122); outputs are the same canonical residues.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint32_t u32;
typedef uint64_t u64;

#define CM_HD __host__ __device__ __forceinline__

// r in [0, 2p) -> [0, p).  r - p wraps above 2^31 when r < p (p < 2^31).
CM_HD u32 cm_reduce2p(u32 r, u32 p) { u32 t = r - p; return t < r ? t : r; }

CM_HD u32 cm_add(u32 a, u32 b, u32 p) { return cm_reduce2p(a + b, p); }

// a - b for canonical a, b: when a < b the difference wraps above 2^31 and
// adding p wraps it back into [0, p), so the smaller of d and d + p is the
// residue (three instructions).
CM_HD u32 cm_sub(u32 a, u32 b, u32 p) {
    const u32 d = a - b, e = d + p;
    return e < d ? e : d;
}

CM_HD u32 cm_mulhi(u32 a, u32 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (u32)(((u64)a * b) >> 32);
#endif
}

// Shoup: a * w mod p, result in [0, 2p).  a may be any 32-bit value.
CM_HD u32 cm_shoup_lazy(u32 a, u32 w, u32 wq, u32 p) {
    u32 q = cm_mulhi(a, wq);
    return a * w - q * p;
}

CM_HD u32 cm_shoup(u32 a, u32 w, u32 wq, u32 p) {
    return cm_reduce2p(cm_shoup_lazy(a, w, wq, p), p);
}

// Lazy Montgomery product a * b / 2^32 mod p in [0, 2p), for a * b < p 2^32
// (any 32-bit a when b < p).  npinv = -p^-1 mod 2^32.  p < 2^31 keeps
// t + m p below 2^64.
CM_HD u32 cm_mont_lazy(u32 a, u32 b, u32 p, u32 npinv) {
    const u64 t = (u64)a * b;
    const u32 m = (u32)t * npinv;
    return (u32)(((u64)m * p + t) >> 32);
}

// a * w mod p, canonical, for a twiddle or constant in Montgomery form wm = w R mod p
CM_HD u32 cm_mul_tw(u32 a, u32 wm, u32 p, u32 npinv) {
    return cm_reduce2p(cm_mont_lazy(a, wm, p, npinv), p);
}

// Montgomery: a * b * 2^-32 mod p for a, b < p.
CM_HD u32 cm_mont(u32 a, u32 b, u32 p, u32 pinv) {
    u64 t = (u64)a * b;
    u32 m = (u32)t * pinv;
    u32 mh = cm_mulhi(m, p);
    u32 r = (u32)(t >> 32) - mh;       // (t - m p) / 2^32 in (-p, p), wrapped
    u32 rp = r + p;                    // wraps back into [0, p) when r < 0
    return rp < r ? rp : r;
}

// ---- host helpers ----------------------------------------------------------
static inline u32 cm_host_mul(u32 a, u32 b, u32 p) { return (u32)(((u64)a * b) % p); }

static inline u32 cm_host_pow(u32 a, u64 e, u32 p) {
    u64 r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (u32)r;
}

static inline u32 cm_host_inv(u32 a, u32 p) { return cm_host_pow(a, (u64)p - 2, p); }

// floor(w * 2^32 / p), the Shoup companion of w (w < p)
static inline u32 cm_host_shoup_q(u32 w, u32 p) { return (u32)(((u64)w << 32) / p); }

// p^-1 mod 2^32 by Newton iteration (p odd)
static inline u32 cm_host_pinv(u32 p) {
    u32 x = p;                        // correct to 3 bits
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return x;
}

// 2^32 mod p
static inline u32 cm_host_r1(u32 p) { return (u32)(((u64)1 << 32) % p); }

// Montgomery form w 2^32 mod p of w < p
static inline u32 cm_host_mont(u32 w, u32 p) { return (u32)(((u64)w << 32) % p); }

// -p^-1 mod 2^32
static inline u32 cm_host_npinv(u32 p) { return 0u - cm_host_pinv(p); }

static inline int cm_fourier_degree(u32 p) {
    u32 q = p - 1;
    int e = 0;
    while (q && !(q & 1)) { q >>= 1; ++e; }
    return e;
}

static inline int cm_is_prime(u32 n) {
    if (n < 2) return 0;
    if (n % 2 == 0) return n == 2;
    for (u32 d = 3; (u64)d * d <= n; d += 2)
        if (n % d == 0) return 0;
    return 1;
}

// smallest generator of (Z/pZ)^*
static inline u32 cm_host_generator(u32 p) {
    u32 fac[32];
    int nf = 0;
    u32 q = p - 1;
    for (u32 d = 2; (u64)d * d <= q; ++d) {
        if (q % d == 0) {
            fac[nf++] = d;
            while (q % d == 0) q /= d;
        }
    }
    if (q > 1) fac[nf++] = q;
    for (u32 g = 2; g < p; ++g) {
        int ok = 1;
        for (int i = 0; i < nf && ok; ++i)
            if (cm_host_pow(g, (p - 1) / fac[i], p) == 1) ok = 0;
        if (ok) return g;
    }
    return 0;
}
