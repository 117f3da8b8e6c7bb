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

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Limb split: coefficient i of a (w words, two's complement) -> rows j < K of
// out (row length Ly): |a_i| bits [jM, jM+M) mod p, negated mod p when
// a_i < 0; rows K .. 2K-1 and positions i >= n are zero.
__global__ void k_tc_split(const u64 *__restrict__ a, int64_t n, int w, int K, int K2, int M,
                           u32 p, u32 *__restrict__ out, int64_t Ly) {
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
                    out[j * Ly + i] = neg ? (limb ? p - limb : 0u) : limb;
                    buf >>= M;
                    have -= M;
                    ++j;
                }
            }
            for (; j < K; ++j) {                 // remaining bits, then zero limbs
                const u32 limb = (u32)((u64)buf & mask);
                out[j * Ly + i] = neg ? (limb ? p - limb : 0u) : limb;
                buf >>= M;
            }
        }
        for (; j < K2; ++j) out[j * Ly + i] = 0u;
    }
}

// CRT and recombination: output coefficient y = sum_j z_j 2^(M j), where z_j
// (j < 2K-1) is the signed CRT value of the residues C_r[j][y].
template <int NP>
__global__ void k_tc_crt(const u32 *__restrict__ C0, const u32 *__restrict__ C1,
                         const u32 *__restrict__ C2, int K2, int M, int64_t Ly, int64_t ny,
                         u64 *__restrict__ out, int wc, u32 p0, u32 p1, u32 p2,
                         u32 inv01, u32 inv01q, u32 inv012, u32 inv012q, u32 p0m2,
                         u32 p0m2q) {
    for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < ny;
         y += (int64_t)gridDim.x * blockDim.x) {
        u64 *o = out + y * wc;
        for (int q = 0; q < wc; ++q) o[q] = 0;
        for (int j = 0; j < K2 - 1; ++j) {
            const u32 r0 = C0[j * Ly + y], r1 = C1[j * Ly + y];
            // Garner: t1 = (r1 - r0) p0^-1 mod p1;  z = r0 + p0 t1 (+ p0 p1 t2)
            const u32 d1 = cm_sub(r1, r0 % p1, p1);
            const u32 t1 = cm_shoup(d1, inv01, inv01q, p1);
            unsigned __int128 z = (unsigned __int128)r0 + (unsigned __int128)p0 * t1;
            unsigned __int128 P = (unsigned __int128)p0 * p1;
            if (NP == 3) {
                const u32 r2 = C2[j * Ly + y];
                // t2 = (r2 - (r0 + p0 t1)) (p0 p1)^-1 mod p2
                const u32 zm = (u32)(((u64)(r0 % p2) + cm_shoup(t1 % p2, p0m2, p0m2q, p2)) % p2);
                const u32 d2 = cm_sub(r2, zm, p2);
                const u32 t2 = cm_shoup(d2, inv012, inv012q, p2);
                z += P * t2;
                P *= p2;
            }
            // signed value in (-P/2, P/2]
            __int128 sz = (__int128)z;
            if (z > P / 2) sz -= (__int128)P;
            if (sz == 0) continue;
            // add sz << (M j) into the wc-word accumulator
            const int sh = M * j, w0 = sh >> 6, b0 = sh & 63;
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
    const int K2 = 1 << pl.lk2;
    const int64_t Ly = (int64_t)1 << pl.ey;
    const int64_t ny = na + nb - 1;
    const size_t per = (size_t)2 * K2 * Ly;   // A and B rows of one prime
    u32 *ws = (u32 *)workspace(10, sizeof(u32) * per * pl.np);
    if (!ws) return hipErrorOutOfMemory;
    u32 *Cr[3] = {nullptr, nullptr, nullptr};
    for (int r = 0; r < pl.np; ++r) {
        const u32 p = kTcPrimes[r];
        const Family *fam = family_default(p);
        if (!fam) return hipErrorOutOfMemory;
        u32 *A = ws + per * r, *B = A + (size_t)K2 * Ly;
        Cr[r] = A;
        {
            ProfScope ps(kProfTcSplit, s);
            hipLaunchKernelGGL(k_tc_split, dim3(blocks_for(Ly)), dim3(256), 0, s, a, na, wa, pl.K,
                               K2, pl.M, p, A, Ly);
            hipLaunchKernelGGL(k_tc_split, dim3(blocks_for(Ly)), dim3(256), 0, s, b, nb, wb, pl.K,
                               K2, pl.M, p, B, Ly);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // x: length-2K DFT over the rows (cyclic points even, negacyclic odd)
        const u32 wx = cm_host_pow(fam->wN, (u64)1 << (fam->K - pl.lk2), p);
        const Family *fx = family_for_root(p, wx, pl.lk2);
        const Family *fxi = family_for_root(p, cm_host_inv(wx, p), pl.lk2);
        if (!fx || !fxi) return hipErrorOutOfMemory;
        e = launch_col_ntt(pl.lk2, A, Ly, 1u, fx, s);
        if (e == hipSuccess) e = launch_col_ntt(pl.lk2, B, Ly, 1u, fx, s);
        // y: 2K independent cyclic products of length Ly (= linear: ny <= Ly)
        if (e == hipSuccess) e = launch_rows_mul(pl.ey, A, B, K2, fam, s);
        // inverse x transform with the 1/(2K) scaling
        const u32 ki = cm_host_inv((u32)K2, p);
        if (e == hipSuccess)
            e = launch_col_ntt(pl.lk2, A, Ly, ki, fxi, s);
        if (e != hipSuccess) return e;
    }
    const u32 p0 = kTcPrimes[0], p1 = kTcPrimes[1], p2 = kTcPrimes[2];
    const u32 inv01 = cm_host_inv(p0 % p1, p1);
    const u32 p0m2 = p0 % p2;
    const u32 inv012 = cm_host_inv(cm_host_mul(p0 % p2, p1 % p2, p2), p2);
    ProfScope ps(kProfTcCrt, s);
    if (pl.np == 2)
        hipLaunchKernelGGL((k_tc_crt<2>), dim3(blocks_for(ny)), dim3(256), 0, s, Cr[0], Cr[1],
                           Cr[1], K2, pl.M, Ly, ny, c, wc, p0, p1, p2, inv01,
                           cm_host_shoup_q(inv01, p1), inv012, cm_host_shoup_q(inv012, p2), p0m2,
                           cm_host_shoup_q(p0m2, p2));
    else
        hipLaunchKernelGGL((k_tc_crt<3>), dim3(blocks_for(ny)), dim3(256), 0, s, Cr[0], Cr[1],
                           Cr[2], K2, pl.M, Ly, ny, c, wc, p0, p1, p2, inv01,
                           cm_host_shoup_q(inv01, p1), inv012, cm_host_shoup_q(inv012, p2), p0m2,
                           cm_host_shoup_q(p0m2, p2));
    return hipGetLastError();
}

hipError_t launch_scale_x(u32 *A, int64_t len, int K, u32 theta, u32 p, hipStream_t s) {
    if (len <= 0 || K <= 0) return hipSuccess;
    std::vector<uint2> tp((size_t)K);
    u64 x = 1;
    for (int j = 0; j < K; ++j) {
        tp[(size_t)j] = make_uint2((u32)x, cm_host_shoup_q((u32)x, p));
        x = x * (theta % p) % p;
    }
    uint2 *d = (uint2 *)workspace(11, sizeof(uint2) * (size_t)K);
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
