// twoconv.hip -- exact product of integer polynomials by two convolutions
// over 32-bit Fourier primes, CRT and carry recombination.
//
// Reference: two_convolution_poly_mul (src/two_convolution_poly_mul.cu:532-569)
// and its parts :58-428.  Each coefficient (sign + magnitude) is cut into K
// limbs of M bits, giving a bivariate A(x, y) with a(y) = A(2^M, y); a negative
// coefficient contributes (p - limb) mod p (:84-203).  The product C = A B has
// x-degree <= 2K-2, and
//     cyclic      C mod (x^K - 1) = C_lo + C_hi
//     negacyclic  C mod (x^K + 1) = C_lo - C_hi   (theta twist, :490-512)
// so c(y) = C_lo(2^M, y) + 2^(KM) C_hi(2^M, y)  (:407-423 recombine the
// CRT-reconstructed sum 2 C_lo and difference 2 C_hi the same way).
//
// Here both convolutions are evaluated at once: the length-2K DFT in x of the
// zero-padded limb vector has the cyclic convolution's points at even indices
// (w_2K^(2v) = w_K^v) and the theta-twisted negacyclic ones at odd indices
// (w_2K^(2v+1) = theta w_K^v, theta = w_2K).  Every x-point is an independent
// linear convolution in y (a row multiply), and the inverse length-2K DFT
// returns C_lo (j < K) and C_hi (j >= K) directly, i.e. the reference's
// (sum, diff)/2.  The reference pads x and y to one square power of two
// (:446-453) and therefore only recovers C_lo; its answers are exact while
// the limb product does not wrap.  Here x has exactly 2K points and y is
// padded to the product length, and (K, M, prime count) are chosen from the
// CRT bound  L * n_min * (2^M - 1)^2 < P / 2  (L = nonzero limb pairs), so the
// result is the exact integer product for every input size.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "ntt_core.h"

namespace cm {

// CRT primes (p - 1 divisible by 2^25 or more) and the inverses Garner needs
static const u32 kTcPrimes[3] = {2013265921u, 2113929217u, 1811939329u};

struct TcPlan {
    int np;        // primes (2 or 3)
    int K, M;      // limbs per coefficient, bits per limb
    int lk2;       // log2(2K)
    int ey;        // y transform size 2^ey >= na + nb - 1
};

// (K, M, np) minimising np * K under the CRT bound; false if none fits.
static bool tc_choose(int bits, int64_t na, int64_t nb, TcPlan &pl) {
    const double lg_nmin = std::log2((double)std::min(na, nb));
    const double lgP[2] = {std::log2(2013265921.0) + std::log2(2113929217.0),
                           std::log2(2013265921.0) + std::log2(2113929217.0) +
                               std::log2(1811939329.0)};
    int best = 0;
    for (int np = 2; np <= 3; ++np) {
        for (int K = 2; K <= 512; K *= 2) {
            const int M = (bits + K - 1) / K;
            if (M > 30 || M < 1) continue;
            const int L = std::min(K, (bits + M - 1) / M);
            const double need = std::log2((double)L) + lg_nmin + 2.0 * M;
            if (need >= lgP[np - 2] - 1.01) continue;
            const int cost = np * K;
            if (!best || cost < best) {
                best = cost;
                pl.np = np; pl.K = K; pl.M = M;
            }
            break;   // larger K only costs more for this np
        }
    }
    if (!best) return false;
    pl.lk2 = 0;
    while ((1 << pl.lk2) < 2 * pl.K) ++pl.lk2;
    pl.ey = 0;
    while (((int64_t)1 << pl.ey) < na + nb - 1) ++pl.ey;
    if (pl.ey < 4) pl.ey = 4;
    return true;
}

// Plan shared by a class of batched products with y size 2^ey: the cheapest
// (np, K) whose widest admissible limb M_max (the CRT bound with every one of
// the K limbs nonzero and n_min = 2^(ey-1), the class's worst case) covers
// `bits`, with M = M_max.  Products of different bit widths then share a
// plan, and so a launch sequence, whenever one (np, K) serves them.
static bool tc_class_plan(int bits, int ey, TcPlan &pl) {
    const double lgP[2] = {std::log2(2013265921.0) + std::log2(2113929217.0),
                           std::log2(2013265921.0) + std::log2(2113929217.0) +
                               std::log2(1811939329.0)};
    int best = 0;
    for (int np = 2; np <= 3; ++np) {
        for (int K = 1; K <= 512; K *= 2) {
            int M = 30;
            while (M >= 1 && std::log2((double)K) + (ey - 1) + 2.0 * M >= lgP[np - 2] - 1.01) --M;
            if (M < 1 || (int64_t)K * M < bits) continue;
            const int cost = np * K;
            if (!best || cost < best) {
                best = cost;
                pl.np = np; pl.K = K; pl.M = M;
            }
            break;
        }
    }
    if (!best) return false;
    pl.lk2 = 0;
    while ((1 << pl.lk2) < 2 * pl.K) ++pl.lk2;
    pl.ey = ey;
    return true;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// One product of a batch (word offsets into A, B, C; coefficient counts;
// words per coefficient).
struct TcDesc {
    int64_t a_off, b_off, c_off, na, nb;
    int32_t wa, wb, wc, pad;
};

// Limb split of one operand of a batch: coefficient i of item r (w words,
// two's complement) -> out[(j R + r) Ly + i] for limb j < K: |a_i| bits
// [jM, jM+M) mod p, negated mod p when a_i < 0; limbs K .. 2K-1 and
// positions i >= n are zero.  grid.y: items (descriptor `one` when d is null).
__global__ void k_tc_split(const u64 *__restrict__ src, const TcDesc *__restrict__ d,
                           TcDesc one, int which, int K, int K2, int M, u32 p,
                           u32 *__restrict__ out, int64_t Ly, int R) {
    const int r = blockIdx.y;
    const TcDesc dd = d ? d[r] : one;
    const u64 *a = src + (which ? dd.b_off : dd.a_off);
    const int64_t n = which ? dd.nb : dd.na;
    const int w = which ? dd.wb : dd.wa;
    u32 *o = out + (int64_t)r * Ly;
    const int64_t js = (int64_t)R * Ly;           // stride between limb rows
    const u64 mask = (1ull << M) - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < Ly;
         i += (int64_t)gridDim.x * blockDim.x) {
        int j = 0;
        if (i < n) {
            const u64 *x = a + i * w;
            const bool neg = (int64_t)x[w - 1] < 0;
            u64 carry = 1;                       // magnitude = ~x + 1 when negative
            unsigned __int128 buf = 0;           // holds `have` < M + 64 bits
            int have = 0;
            for (int q = 0; q < w && j < K; ++q) {
                u64 v = x[q];
                if (neg) {
                    v = ~v + carry;
                    carry = (carry && v == 0) ? 1u : 0u;
                }
                buf |= (unsigned __int128)v << have;
                have += 64;
                while (have >= M && j < K) {
                    const u32 limb = (u32)((u64)buf & mask);
                    o[j * js + i] = neg ? (limb ? p - limb : 0u) : limb;
                    buf >>= M;
                    have -= M;
                    ++j;
                }
            }
            for (; j < K; ++j) {                 // remaining bits, then zero limbs
                const u32 limb = (u32)((u64)buf & mask);
                o[j * js + i] = neg ? (limb ? p - limb : 0u) : limb;
                buf >>= M;
            }
        }
        for (; j < K2; ++j) o[j * js + i] = 0u;
    }
}

// k_tc_split for one-word coefficients fused with the forward x transform,
// for both operands and every prime in one launch (grid.z = 2 q + which):
// thread i forms the K = 2^LK2 / 2 limbs of coefficient i of operand
// `which` (upper K2 - K limbs zero), runs the natural-order length-K2 DFT
// on them in registers (DIF stages with prime q's x-root level table, zero
// upper half) and writes out[q][which][(k R + r) Ly + i] = DFT_k for every
// k < K2.  Positions i >= n are zero rows.
struct TcFwdX {
    const u32 *lv[3];   // x root's level tables, per prime
    u32 p[3], np[3];
    u32 *out[3][2];     // A and B arrays of each prime
};

template <int LK2>
__global__ void k_tc_split_x(const u64 *__restrict__ A, const u64 *__restrict__ B,
                             const TcDesc *__restrict__ d, TcDesc one, int M, TcFwdX fx,
                             int64_t Ly, int R) {
    constexpr int K2 = 1 << LK2, K = K2 / 2;
    const int r = blockIdx.y, q = blockIdx.z >> 1, which = blockIdx.z & 1;
    const TcDesc dd = d ? d[r] : one;
    const u64 *a = which ? B + dd.b_off : A + dd.a_off;
    const int64_t n = which ? dd.nb : dd.na;
    u32 *o = fx.out[q][which] + (int64_t)r * Ly;
    const int64_t js = (int64_t)R * Ly;
    const u64 mask = (1ull << M) - 1;
    const Mod md{fx.p[q], fx.np[q]};
    const u32 *lvx = fx.lv[q];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < Ly;
         i += (int64_t)gridDim.x * blockDim.x) {
        u32 v[1][K2];
#pragma unroll
        for (int j = 0; j < K2; ++j) v[0][j] = 0u;
        if (i < n) {
            const u64 x = a[i];
            const bool neg = (int64_t)x < 0;
            const u64 mag = neg ? ~x + 1 : x;            // |x| (2^63 for the minimum)
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int sh = j * M;
                const u32 limb = sh < 64 ? (u32)((mag >> sh) & mask) : 0u;
                v[0][j] = neg ? (limb ? md.p - limb : 0u) : limb;
            }
        }
        cm_dif_stages<LK2, LK2, 0, 0, 1, true>(v, 0, lvx, md);
#pragma unroll
        for (int m = 0; m < K2; ++m) o[(int64_t)cm_brev((u32)m, LK2) * js + i] = v[0][m];
    }
}

// Garner constants of the CRT over the first two or three primes
struct TcCrt {
    u32 p0, p1, p2, inv01, inv01q, inv012, inv012q, p0m2, p0m2q;
};

// Adds z 2^sh to the wc-word two's-complement accumulator o, where z is the
// signed CRT value (in (-P/2, P/2]) of the residues r0, r1 (, r2).
template <int NP>
__device__ __forceinline__ void tc_crt_add(u64 *o, int wc, int sh, u32 r0, u32 r1, u32 r2,
                                           const TcCrt &k) {
    // Garner: t1 = (r1 - r0) p0^-1 mod p1;  z = r0 + p0 t1 (+ p0 p1 t2)
    const u32 d1 = cm_sub(r1, r0 % k.p1, k.p1);
    const u32 t1 = cm_shoup(d1, k.inv01, k.inv01q, k.p1);
    unsigned __int128 z = (unsigned __int128)r0 + (unsigned __int128)k.p0 * t1;
    unsigned __int128 P = (unsigned __int128)k.p0 * k.p1;
    if (NP == 3) {
        // t2 = (r2 - (r0 + p0 t1)) (p0 p1)^-1 mod p2
        const u32 zm =
            (u32)(((u64)(r0 % k.p2) + cm_shoup(t1 % k.p2, k.p0m2, k.p0m2q, k.p2)) % k.p2);
        const u32 d2 = cm_sub(r2, zm, k.p2);
        const u32 t2 = cm_shoup(d2, k.inv012, k.inv012q, k.p2);
        z += P * t2;
        P *= k.p2;
    }
    // signed value in (-P/2, P/2]
    __int128 sz = (__int128)z;
    if (z > P / 2) sz -= (__int128)P;
    if (sz == 0) return;
    // add sz << sh into the wc-word accumulator
    const int w0 = sh >> 6, b0 = sh & 63;
    const bool neg = sz < 0;
    unsigned __int128 mag = (unsigned __int128)(neg ? -sz : sz);
    u64 part[3];
    part[0] = (u64)(mag << b0);
    part[1] = (u64)((mag << b0) >> 64);
    part[2] = b0 ? (u64)(mag >> (128 - b0)) : 0;
    if (!neg) {
        u64 c = 0;
        for (int q = w0; q < wc; ++q) {
            const u64 s = (q - w0 < 3) ? part[q - w0] : 0;
            const u64 t = o[q] + s;
            const u64 c1 = t < s;
            const u64 t2 = t + c;
            c = c1 | (t2 < c);
            o[q] = t2;
            if (q - w0 >= 2 && !c) break;
        }
    } else {
        u64 b = 0;
        for (int q = w0; q < wc; ++q) {
            const u64 s = (q - w0 < 3) ? part[q - w0] : 0;
            const u64 v = o[q];
            const u64 t = v - s - b;
            b = (v < s) || (v - s < b);
            o[q] = t;
            if (q - w0 >= 2 && !b) break;
        }
    }
}

// CRT and recombination for a batch: output coefficient y of item r is
// sum_j z_j 2^(M j), where z_j (j < 2K-1) is the signed CRT value of the
// residues C_q[(j R + r) Ly + y].  grid.y: items.
template <int NP>
__global__ void k_tc_crt(const u32 *__restrict__ C0, const u32 *__restrict__ C1,
                         const u32 *__restrict__ C2, const TcDesc *__restrict__ d, TcDesc one,
                         int K2, int M, int64_t Ly, int R, u64 *__restrict__ out, TcCrt k) {
    const int r = blockIdx.y;
    const TcDesc dd = d ? d[r] : one;
    const int64_t ny = dd.na + dd.nb - 1;
    const int wc = dd.wc;
    const int64_t js = (int64_t)R * Ly, base = (int64_t)r * Ly;
    for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < ny;
         y += (int64_t)gridDim.x * blockDim.x) {
        u64 *o = out + dd.c_off + y * wc;
        for (int q = 0; q < wc; ++q) o[q] = 0;
        for (int j = 0; j < K2 - 1; ++j) {
            const int64_t at = j * js + base + y;
            tc_crt_add<NP>(o, wc, M * j, C0[at], C1[at], NP == 3 ? C2[at] : 0u, k);
        }
    }
}

// Inverse x transforms of the residues of each prime (natural-order length-K2
// DFT with the inverse root's level table, times 1/K2) fused with the CRT
// and recombination: thread y loads C_q[(k R + r) Ly + y] for k < K2 and
// prime q, transforms them in registers, then adds z_j 2^(M j) for
// j < K2 - 1 as k_tc_crt does.
struct TcInvX {
    const u32 *lv[3];   // inverse x root's level tables, per prime
    u32 np[3];          // -p^-1 mod 2^32, per prime
    u32 ki[3];          // Montgomery form of 1 / K2, per prime
};

template <int NP, int LK2>
__global__ void k_tc_crt_x(const u32 *__restrict__ C0, const u32 *__restrict__ C1,
                           const u32 *__restrict__ C2, const TcDesc *__restrict__ d, TcDesc one,
                           int M, int64_t Ly, int R, u64 *__restrict__ out, TcCrt k, TcInvX ix) {
    constexpr int K2 = 1 << LK2;
    const int r = blockIdx.y;
    const TcDesc dd = d ? d[r] : one;
    const int64_t ny = dd.na + dd.nb - 1;
    const int wc = dd.wc;
    const int64_t js = (int64_t)R * Ly, base = (int64_t)r * Ly;
    const u32 ps[3] = {k.p0, k.p1, k.p2};
    const u32 *Cs[3] = {C0, C1, C2};
    for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < ny;
         y += (int64_t)gridDim.x * blockDim.x) {
        u32 res[NP][K2];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            u32 v[1][K2];
#pragma unroll
            for (int j = 0; j < K2; ++j) v[0][j] = Cs[q][j * js + base + y];
            const Mod md{ps[q], ix.np[q]};
            cm_dif_stages<LK2, LK2, 0, 0, 1>(v, 0, ix.lv[q], md);
#pragma unroll
            for (int j = 0; j < K2; ++j)
                res[q][j] = cm_mul_tw(v[0][cm_brev((u32)j, LK2)], ix.ki[q], md.p, md.np);
        }
        u64 *o = out + dd.c_off + y * wc;
        for (int q = 0; q < wc; ++q) o[q] = 0;
#pragma unroll
        for (int j = 0; j < K2 - 1; ++j)
            tc_crt_add<NP>(o, wc, M * j, res[0][j], res[1][j], NP == 3 ? res[2][j] : 0u, k);
    }
}

// theta scaling of the reference's scale_x_argument_dev
// (src/two_convolution_poly_mul.cu:207-245): A[i] *= theta^(i mod K)
__global__ void k_scale_x(u32 *__restrict__ A, int64_t len, int K, const uint2 *__restrict__ tp,
                          u32 p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint2 t = tp[i % K];
        A[i] = cm_shoup(A[i] % p, t.x, t.y, p);
    }
}

static int blocks_for(int64_t n, int t = 256) {
    int64_t b = (n + t - 1) / t;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
struct TcWorkspace {
    u32 *rows = nullptr;   // per prime: A (2K x Ly) then B (2K x Ly)
};

int tc_bit_width(const u64 *a, int64_t n, int w) {
    int best = 1;
    std::vector<u64> mag((size_t)w);
    for (int64_t i = 0; i < n; ++i) {
        const u64 *x = a + i * w;
        const bool neg = (int64_t)x[w - 1] < 0;
        u64 carry = 1;
        for (int q = 0; q < w; ++q) {
            u64 v = x[q];
            if (neg) { v = ~v + carry; carry = (carry && v == 0); }
            mag[(size_t)q] = v;
        }
        int l = w;
        while (l > 0 && mag[(size_t)l - 1] == 0) --l;
        if (l) {
            int bits = 64 * (l - 1) + (64 - __builtin_clzll(mag[(size_t)l - 1]));
            best = std::max(best, bits);
        }
    }
    return best;
}

// The products of one class (same plan and y size) over R items: per prime,
// A and B are 2K x R x Ly (limb row j of item r at (j R + r) Ly), so the x
// transforms are column transforms over R Ly columns and the y products are
// 2K R rows of Ly words.  d: device descriptors, or null with `one`.  w1:
// every coefficient of the class is one word; then, for 2K <= 8, the x
// transforms run inside the split and CRT kernels (k_tc_split_x,
// k_tc_crt_x) instead of as three column passes per prime.
static hipError_t tc_run_class(const TcPlan &pl, const u64 *A, const u64 *B, u64 *C,
                               const TcDesc *d, const TcDesc &one, int R, int64_t ymax, bool w1,
                               u32 *ws, hipStream_t s) {
    const int K2 = 1 << pl.lk2;
    const int64_t Ly = (int64_t)1 << pl.ey;
    const size_t per = (size_t)2 * K2 * R * Ly;   // A and B rows of one prime
    const int bx = blocks_for(Ly);
    const bool fused = w1 && pl.lk2 >= 1 && pl.lk2 <= 3;   // 2K = 16 would spill in the CRT
    u32 *Cr[3] = {nullptr, nullptr, nullptr};
    TcInvX ix{};
    TcFwdX fw{};
    const Family *fams[3] = {nullptr, nullptr, nullptr};
    for (int q = 0; q < pl.np; ++q) {
        const u32 p = kTcPrimes[q];
        const Family *fam = family_default(p);
        if (!fam) return hipErrorOutOfMemory;
        fams[q] = fam;
        u32 *Aq = ws + per * q, *Bq = Aq + (size_t)K2 * R * Ly;
        Cr[q] = Aq;
        // x: length-2K DFT down the columns (cyclic points even, negacyclic odd)
        const u32 wx = cm_host_pow(fam->wN, (u64)1 << (fam->K - pl.lk2), p);
        const Family *fx = family_for_root(p, wx, pl.lk2);
        const Family *fxi = family_for_root(p, cm_host_inv(wx, p), pl.lk2);
        if (!fx || !fxi) return hipErrorOutOfMemory;
        const u32 ki = cm_host_inv((u32)K2, p);
        if (fused) {
            fw.lv[q] = fx->lv;
            fw.p[q] = p;
            fw.np[q] = fam->np;
            fw.out[q][0] = Aq;
            fw.out[q][1] = Bq;
            ix.lv[q] = fxi->lv;
            ix.np[q] = fam->np;
            ix.ki[q] = cm_host_mont(ki, p);
            continue;
        }
        hipError_t e;
        {
            ProfScope ps(kProfTcSplit, s);
            hipLaunchKernelGGL(k_tc_split, dim3(bx, R), dim3(256), 0, s, A, d, one, 0, pl.K, K2,
                               pl.M, p, Aq, Ly, R);
            hipLaunchKernelGGL(k_tc_split, dim3(bx, R), dim3(256), 0, s, B, d, one, 1, pl.K, K2,
                               pl.M, p, Bq, Ly, R);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = launch_col_ntt(pl.lk2, Aq, (int64_t)R * Ly, 1u, fx, s);
        if (e == hipSuccess) e = launch_col_ntt(pl.lk2, Bq, (int64_t)R * Ly, 1u, fx, s);
        // y: 2K R independent cyclic products of length Ly (= linear: ny <= Ly)
        if (e == hipSuccess) e = launch_rows_mul(pl.ey, Aq, Bq, K2 * R, fam, s);
        // inverse x transform with the 1/(2K) scaling
        if (e == hipSuccess) e = launch_col_ntt(pl.lk2, Aq, (int64_t)R * Ly, ki, fxi, s);
        if (e != hipSuccess) return e;
    }
    if (fused) {
        // both operands of every prime split and transformed in one launch,
        // then the y products prime by prime
        {
            ProfScope ps(kProfTcSplit, s);
            const dim3 g(bx, R, 2 * pl.np);
            if (pl.lk2 == 1)
                hipLaunchKernelGGL(k_tc_split_x<1>, g, dim3(256), 0, s, A, B, d, one, pl.M, fw, Ly, R);
            else if (pl.lk2 == 2)
                hipLaunchKernelGGL(k_tc_split_x<2>, g, dim3(256), 0, s, A, B, d, one, pl.M, fw, Ly, R);
            else
                hipLaunchKernelGGL(k_tc_split_x<3>, g, dim3(256), 0, s, A, B, d, one, pl.M, fw, Ly, R);
        }
        hipError_t e = hipGetLastError();
        for (int q = 0; q < pl.np && e == hipSuccess; ++q)
            e = launch_rows_mul(pl.ey, fw.out[q][0], fw.out[q][1], K2 * R, fams[q], s);
        if (e != hipSuccess) return e;
    }
    const u32 p0 = kTcPrimes[0], p1 = kTcPrimes[1], p2 = kTcPrimes[2];
    const u32 inv01 = cm_host_inv(p0 % p1, p1);
    const u32 p0m2 = p0 % p2;
    const u32 inv012 = cm_host_inv(cm_host_mul(p0 % p2, p1 % p2, p2), p2);
    const TcCrt kc{p0, p1, p2, inv01, cm_host_shoup_q(inv01, p1), inv012,
                   cm_host_shoup_q(inv012, p2), p0m2, cm_host_shoup_q(p0m2, p2)};
    ProfScope ps(kProfTcCrt, s);
    const dim3 g(blocks_for(ymax), R);
    const u32 *C2 = pl.np == 3 ? Cr[2] : Cr[1];
    if (fused) {
#define CM_CX(NP, L)                                                                        \
    hipLaunchKernelGGL((k_tc_crt_x<NP, L>), g, dim3(256), 0, s, Cr[0], Cr[1], C2, d, one, pl.M, \
                       Ly, R, C, kc, ix)
        if (pl.np == 2) {
            switch (pl.lk2) {
            case 1: CM_CX(2, 1); break;
            case 2: CM_CX(2, 2); break;
            default: CM_CX(2, 3); break;
            }
        } else {
            switch (pl.lk2) {
            case 1: CM_CX(3, 1); break;
            case 2: CM_CX(3, 2); break;
            default: CM_CX(3, 3); break;
            }
        }
#undef CM_CX
    } else if (pl.np == 2) {
        hipLaunchKernelGGL((k_tc_crt<2>), g, dim3(256), 0, s, Cr[0], Cr[1], C2, d, one, K2, pl.M,
                           Ly, R, C, kc);
    } else {
        hipLaunchKernelGGL((k_tc_crt<3>), g, dim3(256), 0, s, Cr[0], Cr[1], C2, d, one, K2, pl.M,
                           Ly, R, C, kc);
    }
    return hipGetLastError();
}

// Runs the whole product on device buffers.  bits: largest coefficient bit
// width of a and b (>= 1).
hipError_t launch_two_convolution(const u64 *a, int64_t na, int wa, const u64 *b, int64_t nb,
                                  int wb, u64 *c, int wc, int bits, hipStream_t s, int *err) {
    *err = 0;
    TcPlan pl;
    if (!tc_choose(bits, na, nb, pl) || pl.ey > kMaxLog || pl.lk2 > kMaxLdsLog) {
        *err = 1;
        return hipSuccess;
    }
    const size_t per = (size_t)2 * (1 << pl.lk2) * ((size_t)1 << pl.ey);
    u32 *ws = (u32 *)workspace(10, sizeof(u32) * per * pl.np, s);
    if (!ws) return hipErrorOutOfMemory;
    TcDesc one{0, 0, 0, na, nb, wa, wb, wc, 0};
    return tc_run_class(pl, a, b, c, nullptr, one, 1, na + nb - 1, wa == 1 && wb == 1, ws, s);
}

// Batch of products: items are grouped by class plan (tc_class_plan: one plan
// per y size and (np, K)) and each class runs as one sequence of batched
// launches, in chunks whose scratch stays under kTcChunkBytes.
static constexpr size_t kTcChunkBytes = (size_t)1 << 30;

// Class merging of the batched multiply: 0 keeps one class per exact plan;
// 1 gives every item of one y size the plan of the widest coefficients of
// that size (exact for all of them: K M covers every width and the bound
// depends on the y size only), trading transform work for fewer launches.
static int g_tc_merge = 0;

bool set_zmul_merge(int m) {
    if (m < 0 || m > 1) return false;
    g_tc_merge = m;
    return true;
}

int zmul_merge() { return g_tc_merge; }

hipError_t launch_two_convolution_list(const TcItem *items, int64_t count, const u64 *A,
                                       const u64 *B, u64 *C, hipStream_t s, int *err) {
    *err = 0;
    struct Cls {
        TcPlan pl;
        std::vector<int64_t> idx;
    };
    std::vector<Cls> classes;
    auto y_size = [](const TcItem &it) {
        int ey = 0;
        while (((int64_t)1 << ey) < it.na + it.nb - 1) ++ey;
        return ey < 4 ? 4 : ey;
    };
    int widest[64] = {0};   // merged classes: widest coefficients per y size
    if (g_tc_merge)
        for (int64_t i = 0; i < count; ++i) {
            const int ey = y_size(items[i]);
            if (ey < 64) widest[ey] = std::max(widest[ey], (int)items[i].bits);
        }
    for (int64_t i = 0; i < count; ++i) {
        const TcItem &it = items[i];
        const int ey = y_size(it);
        const int bits = (g_tc_merge && ey < 64) ? widest[ey] : it.bits;
        TcPlan pl;
        if (!tc_class_plan(bits, ey, pl) || pl.lk2 > kMaxLdsLog) {
            *err = 1;
            return hipSuccess;
        }
        if (pl.ey > kMaxLog) {
            *err = 1;
            return hipSuccess;
        }
        Cls *c = nullptr;
        for (auto &x : classes)
            if (x.pl.np == pl.np && x.pl.K == pl.K && x.pl.M == pl.M && x.pl.ey == pl.ey) c = &x;
        if (!c) {
            classes.push_back(Cls{pl, {}});
            c = &classes.back();
        }
        c->idx.push_back(i);
    }
    // descriptors of every class, in class order, uploaded once
    std::vector<TcDesc> hd((size_t)count);
    size_t at = 0;
    for (auto &c : classes)
        for (int64_t i : c.idx) {
            const TcItem &it = items[i];
            hd[at++] = TcDesc{it.a_off, it.b_off, it.c_off, it.na, it.nb, it.wa, it.wb, it.wc, 0};
        }
    TcDesc *dd = (TcDesc *)workspace(14, sizeof(TcDesc) * hd.size(), s);
    if (!dd) return hipErrorOutOfMemory;
    hipError_t e = stage_upload(hd.data(), sizeof(TcDesc) * hd.size(), dd, s);
    if (e != hipSuccess) return e;
    // Classes are independent: with more than one they run on the caller
    // stream's auxiliary streams (forked after the descriptor upload, joined
    // back at the end), so one class's launch gaps overlap another's work.
    // Each auxiliary stream has its own scratch.
    const int nclass = (int)classes.size();
    AuxStreams *aux = nclass > 1 ? aux_streams(s) : nullptr;
    const int nuse = aux ? std::min(nclass, (int)kAuxStreams) : 0;
    if (aux) {
        e = hipEventRecord(aux->fork, s);
        for (int i = 0; i < nuse && e == hipSuccess; ++i)
            e = hipStreamWaitEvent(aux->s[i], aux->fork, 0);
        if (e != hipSuccess) return e;
    }
    size_t first = 0;
    for (int ci = 0; ci < nclass; ++ci) {
        const auto &c = classes[(size_t)ci];
        const hipStream_t cs = aux ? aux->s[ci % nuse] : s;
        const int K2 = 1 << c.pl.lk2;
        const int64_t Ly = (int64_t)1 << c.pl.ey;
        const size_t item_bytes = sizeof(u32) * (size_t)2 * K2 * Ly * c.pl.np;
        int64_t chunk = (int64_t)(kTcChunkBytes / item_bytes);
        if (chunk < 1) chunk = 1;
        const int64_t n = (int64_t)c.idx.size();
        for (int64_t k0 = 0; k0 < n; k0 += chunk) {
            const int R = (int)std::min<int64_t>(chunk, n - k0);
            u32 *ws = (u32 *)workspace(10, item_bytes * (size_t)R, cs);
            if (!ws) return hipErrorOutOfMemory;
            int64_t ymax = 1;
            bool w1 = true;
            for (int r = 0; r < R; ++r) {
                const TcItem &it = items[c.idx[(size_t)(k0 + r)]];
                ymax = std::max<int64_t>(ymax, it.na + it.nb - 1);
                w1 = w1 && it.wa == 1 && it.wb == 1;
            }
            TcDesc none{};
            e = tc_run_class(c.pl, A, B, C, dd + first + k0, none, R, ymax, w1, ws, cs);
            if (e != hipSuccess) return e;
        }
        first += (size_t)n;
    }
    for (int i = 0; i < nuse; ++i) {
        e = hipEventRecord(aux->join[i], aux->s[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, aux->join[i], 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_scale_x(u32 *A, int64_t len, int K, u32 theta, u32 p, hipStream_t s) {
    if (len <= 0 || K <= 0) return hipSuccess;
    std::vector<uint2> tp((size_t)K);
    u64 x = 1;
    for (int j = 0; j < K; ++j) {
        tp[(size_t)j] = make_uint2((u32)x, cm_host_shoup_q((u32)x, p));
        x = x * (theta % p) % p;
    }
    uint2 *d = (uint2 *)workspace(11, sizeof(uint2) * (size_t)K, s);
    if (!d) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpy(d, tp.data(), sizeof(uint2) * (size_t)K, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scale_x, dim3(blocks_for(len)), dim3(256), 0, s, A, len, K, d, p);
    return hipGetLastError();
}

int tc_plan_info(int bits, int64_t na, int64_t nb, int *np, int *K, int *M) {
    TcPlan pl;
    if (!tc_choose(bits, na, nb, pl)) return 0;
    *np = pl.np; *K = pl.K; *M = pl.M;
    return 1;
}

}  // namespace cm
