// kernels_ntt.hip -- NTT, NTT-multiply and elementwise kernels for gfx950.
//
// Replaces the reference's one-stage-per-launch radix-2 Stockham passes
// (src/stockham.cu:75-397, src/list_stockham.cu:108-576, src/fft_aux.cu:70-322)
// with:
//   * k_mul_lds   -- a whole multiply (forward F and G, pointwise product,
//                    inverse, scaling) inside one workgroup for 2^k <= 2^13;
//   * k_col_fwd / k_row_mul / k_col_inv -- a three-kernel four-step multiply
//                    for larger sizes: column transforms of both operands with
//                    the zero-padding gather fused; row pass with the outer
//                    twiddles, both row transforms, the pointwise product and
//                    the inverse row transform fused; inverse column
//                    transforms with the truncating store fused;
//   * k_ntt_lds / k_row_out / k_col_ntt -- natural-order transforms.
#include <algorithm>

#include "internal.h"
#include "ntt_core.h"

namespace cm {

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so that neighbouring column tiles (which
// share 128-byte lines) meet in the same L2.
CM_DEV int xcd_remap(int b, int nb) {
    if (nb & 7) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// whole multiply in one workgroup (C products per workgroup)
// ---------------------------------------------------------------------------
// DESC: descriptors from `descs` (or the by-value `one` when descs is null);
// otherwise product d reads F + d * one.f, G + d * one.g (one.lf / one.lg
// words) and writes H + d * one.h (one.lh words).  H may alias F.
// Held to 128 VGPRs so that the 2^12- and 2^13-point tiles keep as many
// workgroups per CU as their LDS allows.
template <int LOGL, int C, bool DESC, class MD>
__global__ __launch_bounds__((Tile<LOGL, C, false>::T))
__attribute__((amdgpu_waves_per_eu(4, 8))) void k_mul_lds(
    const u32 *F, const u32 *__restrict__ G, u32 *H, const MulDesc *__restrict__ descs,
    int ndesc, MulDesc one, const u32 *__restrict__ lv, const u32 *__restrict__ lvi, MD md,
    u32 sc) {
    using TL = Tile<LOGL, C, false>;
    __shared__ u32 lds[2 * TL::LDS_WORDS];
    u32 *A = lds, *B = lds + TL::LDS_WORDS;
    const int L = TL::L;

    if constexpr (Round<TL, 0>::R == 4 && Round<TL, 0>::GPT == 1 && TL::NR >= 2) {
        // thread t's product is c = t >> (LOGL-4); it loads its group of the
        // first forward and the last inverse round (elements gi + r 2^(LOGL-4)),
        // which run on the registers
        constexpr int S = LOGL - 4;
        constexpr int PER = TL::WORDS / TL::T;
        const int c = threadIdx.x >> S, gi = threadIdx.x & ((1 << S) - 1);
        const int d = blockIdx.x * C + c;
        MulDesc my{};
        if (d < ndesc) {
            if (DESC) {
                my = descs ? descs[d] : one;
            } else {
                // product d at d times the strides one.f / one.g / one.h
                my = one;
                my.f = (int64_t)d * one.f; my.g = (int64_t)d * one.g; my.h = (int64_t)d * one.h;
            }
        }
        u32 fa[PER], ga[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = gi + (r << S);
            fa[r] = i < my.lf ? F[my.f + i] : 0u;
            ga[r] = i < my.lg ? G[my.g + i] : 0u;
        }
        cm_row_multiply<TL>(A, B, fa, ga, lv, lvi, md, sc);
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = gi + (r << S);
            if (i < my.lh) H[my.h + i] = cm_canon(fa[r], md);
        }
        // words beyond the transform size (lh > L) are zero
        for (int i = L + gi; i < my.lh; i += 1 << S) H[my.h + i] = 0u;
        return;
    } else {

        int64_t off_f[C], off_g[C], off_h[C];
        int lf[C], lg[C], lh[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int d = blockIdx.x * C + c;
            if (DESC) {
                if (d < ndesc) {
                    const MulDesc dd = descs ? descs[d] : one;
                    off_f[c] = dd.f; off_g[c] = dd.g; off_h[c] = dd.h;
                    lf[c] = dd.lf; lg[c] = dd.lg; lh[c] = dd.lh;
                } else {
                    off_f[c] = off_g[c] = off_h[c] = 0;
                    lf[c] = lg[c] = lh[c] = 0;
                }
            } else {
                // product d at d times the strides one.f / one.g / one.h, with
                // the lengths one.lf / one.lg / one.lh
                off_f[c] = (int64_t)d * one.f; off_g[c] = (int64_t)d * one.g;
                off_h[c] = (int64_t)d * one.h;
                lf[c] = d < ndesc ? one.lf : 0; lg[c] = d < ndesc ? one.lg : 0;
                lh[c] = d < ndesc ? one.lh : 0;
            }
        }
        // all loads of a thread are issued before the first LDS store
        constexpr int PER = TL::WORDS / TL::T;
        u32 fa[PER], ga[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL, i = idx & (L - 1);
            fa[r] = i < lf[c] ? F[off_f[c] + i] : 0u;
            ga[r] = i < lg[c] ? G[off_g[c] + i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL, i = idx & (L - 1);
            A[TL::addr(c, i)] = fa[r];
            B[TL::addr(c, i)] = ga[r];
        }
        __syncthreads();
        cm_tile_multiply<TL>(A, B, lv, lvi, md, sc);
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL, i = idx & (L - 1);
            if (i < lh[c]) H[off_h[c] + i] = cm_canon(A[TL::addr(c, i)], md);
        }
        // words beyond the transform size (lh > L) are zero
#pragma unroll
        for (int c = 0; c < C; ++c)
            for (int i = L + threadIdx.x; i < lh[c]; i += TL::T) H[off_h[c] + i] = 0u;
    }
}

// ---------------------------------------------------------------------------
// natural-order forward DFT of rows inside one workgroup
// ---------------------------------------------------------------------------
template <int LOGL, int C>
__global__ __launch_bounds__((Tile<LOGL, C, false>::T)) void k_ntt_lds(
    u32 *__restrict__ X, int64_t m, const u32 *__restrict__ lv, Mod md, u32 sc) {
    using TL = Tile<LOGL, C, false>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const int L = TL::L;
    const int64_t r0 = (int64_t)blockIdx.x * C;
    constexpr int PER = TL::WORDS / TL::T;
    u32 xa[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL, i = idx & (L - 1);
        xa[r] = (r0 + c < m) ? X[(r0 + c) * L + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL, i = idx & (L - 1);
        lds[TL::addr(c, i)] = xa[r];
    }
    __syncthreads();
    cm_tile_dif<TL>(lds, lv, md);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (r0 + c >= m) break;
        for (int i = threadIdx.x; i < L; i += TL::T) {
            u32 v = lds[TL::addr(c, (int)cm_brev(i, LOGL))];
            if (sc) v = cm_mul_tw(v, sc, md.p, md.np);
            X[(r0 + c) * L + i] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// four-step pieces: an n-word array viewed as 2^k1 rows x 2^k2 columns
// ---------------------------------------------------------------------------
// One set of arrays for the column passes: array a reads in[a*in_stride + e]
// for e < in_len (zero beyond) and writes out[a*out_stride + e] for
// e < out_len.
struct ColSet {
    const u32 *in;
    u32 *out;
    int64_t in_stride, in_len, out_stride, out_len;
};

// Column tiles move through memory either one word per lane or, when VEC
// (C a multiple of 4, 16-byte aligned arrays and strides), four consecutive
// columns of a row per lane: a quarter of the memory instructions, each 16
// bytes.  The tile's LDS layout (element (i, c) at TL::phys(i C + c)) is the
// same either way.

// Forward column transforms (DIF, bit-reversed rows), followed by the outer
// twiddle w_n^(i2 * k1) when OUTER (the multiply applies it in the row pass
// instead).  grid.x: column tiles, grid.y: arrays (the first cnt0 from set
// s0, the rest from s1).
template <int LOGL1, int C, bool OUTER, bool VEC, bool TOPZERO = false>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T))
__attribute__((amdgpu_waves_per_eu(8, 8))) void k_col_fwd(
    ColSet s0, ColSet s1, int cnt0, int logn2, const u32 *__restrict__ lv,
    const u32 *__restrict__ olo, const u32 *__restrict__ ohi, Mod md) {
    using TL = Tile<LOGL1, C, true>;
    static_assert(!VEC || (C % 4 == 0 && !OUTER), "vector column access");
    __shared__ __align__(16) u32 lds[TL::LDS_WORDS];
    const int y = blockIdx.y;
    const ColSet &S = y < cnt0 ? s0 : s1;
    const int64_t a = y < cnt0 ? y : y - cnt0;
    const u32 *X = S.in + a * S.in_stride;
    u32 *Y = S.out + a * S.out_stride;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    constexpr int PER = TL::WORDS / TL::T;
    // TOPZERO: every array is shorter than half the transform, so the rows
    // i >= 2^(LOGL1-1) are zero; they are neither read nor stored (the first
    // round treats them as zero)
    constexpr int HALF_ROWS = TL::L / 2;
    if constexpr (VEC) {
        constexpr int PV = PER / 4;
        uint4 xa[PV];
#pragma unroll
        for (int r = 0; r < PV; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T), i = idx / C, c = idx % C;
            const int64_t e = i * n2 + c0 + c;
            if (TOPZERO && i >= HALF_ROWS) continue;
            if (e + 3 < S.in_len) {
                xa[r] = *(const uint4 *)(X + e);
            } else {
                xa[r].x = e < S.in_len ? X[e] : 0u;
                xa[r].y = e + 1 < S.in_len ? X[e + 1] : 0u;
                xa[r].z = e + 2 < S.in_len ? X[e + 2] : 0u;
                xa[r].w = 0u;
            }
        }
#pragma unroll
        for (int r = 0; r < PV; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T);
            if (TOPZERO && idx / C >= HALF_ROWS) continue;
            cm_lds_put4<TL>(lds, idx, xa[r]);
        }
    } else {
        u32 xa[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = threadIdx.x + r * TL::T, i = idx / C, c = idx % C;
            const int64_t e = i * n2 + c0 + c;
            xa[r] = (TOPZERO && i >= HALF_ROWS) ? 0u : (e < S.in_len ? X[e] : 0u);
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            if (TOPZERO && (threadIdx.x + r * TL::T) / C >= HALF_ROWS) continue;
            lds[TL::phys(threadIdx.x + r * TL::T)] = xa[r];
        }
    }
    __syncthreads();
    if constexpr (TOPZERO) cm_tile_dif_topzero<TL>(lds, lv, md);
    else cm_tile_dif<TL>(lds, lv, md);
    if constexpr (VEC) {
#pragma unroll
        for (int r = 0; r < PER / 4; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T), i = idx / C, c = idx % C;
            const uint4 v = cm_lds_get4<TL>(lds, idx);
            *(uint4 *)(Y + i * n2 + c0 + c) = v;
        }
    } else {
        constexpr int LOM = (1 << kOuterLoBits) - 1;
        for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
            const int i = idx / C, c = idx % C;
            u32 v = lds[TL::phys(idx)];
            if (OUTER) {
                const u32 e = (u32)(c0 + c) * cm_brev(i, LOGL1);
                v = cm_mul_tw(cm_mont_lazy(v, olo[e & LOM], md.p, md.np),
                              ohi[e >> kOuterLoBits], md.p, md.np);
            }
            Y[i * n2 + c0 + c] = v;
        }
    }
}

// Inverse column transforms (DIT, bit-reversed rows in, natural rows out)
// with a truncating store; the inverse outer twiddle was applied by the row
// pass.  grid.y: arrays.  Held to 64 VGPRs: at 65 only seven waves fit per
// SIMD, i.e. one 16-wave workgroup per CU instead of two.
template <int LOGL1, int C, bool VEC>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T))
__attribute__((amdgpu_waves_per_eu(8, 8))) void k_col_inv(
    ColSet S, int logn2, const u32 *__restrict__ lvi, Mod md) {
    using TL = Tile<LOGL1, C, true>;
    static_assert(!VEC || C % 4 == 0, "vector column access");
    __shared__ __align__(16) u32 lds[TL::LDS_WORDS];
    const int64_t a = blockIdx.y;
    const u32 *X = S.in + a * S.in_stride;
    u32 *Y = S.out + a * S.out_stride;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    constexpr int PER = TL::WORDS / TL::T;
    if constexpr (VEC) {
        constexpr int PV = PER / 4;
        uint4 xa[PV];
#pragma unroll
        for (int r = 0; r < PV; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T), i = idx / C, c = idx % C;
            xa[r] = *(const uint4 *)(X + i * n2 + c0 + c);
        }
#pragma unroll
        for (int r = 0; r < PV; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T);
            cm_lds_put4<TL>(lds, idx, xa[r]);
        }
    } else {
        u32 xa[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = threadIdx.x + r * TL::T, i = idx / C, c = idx % C;
            xa[r] = X[i * n2 + c0 + c];
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) lds[TL::phys(threadIdx.x + r * TL::T)] = xa[r];
    }
    __syncthreads();
    cm_tile_dit<TL>(lds, lvi, md);
    if constexpr (VEC) {
#pragma unroll
        for (int r = 0; r < PER / 4; ++r) {
            const int idx = 4 * (threadIdx.x + r * TL::T), i = idx / C, c = idx % C;
            const int64_t e = i * n2 + c0 + c;
            const uint4 v = cm_lds_get4<TL>(lds, idx);
            if (e + 3 < S.out_len) {
                *(uint4 *)(Y + e) = v;
            } else {
                if (e < S.out_len) Y[e] = v.x;
                if (e + 1 < S.out_len) Y[e + 1] = v.y;
                if (e + 2 < S.out_len) Y[e + 2] = v.z;
            }
        }
    } else {
        for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
            const int i = idx / C, c = idx % C;
            const int64_t e = i * n2 + c0 + c;
            if (e < S.out_len) Y[e] = lds[TL::phys(idx)];
        }
    }
}

// The same forward and inverse column passes of the multiply with the first
// and last rounds reading and writing global memory directly (no LDS staging
// pass, two barriers fewer; any alignment).  Tiles of two or more rounds.
template <int LOGL1, int C, bool TOPZERO, class MD>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T))
__attribute__((amdgpu_waves_per_eu(8, 8))) void k_col_fwd_d(
    ColSet s0, ColSet s1, int cnt0, int logn2, const u32 *__restrict__ lv, MD md) {
    using TL = Tile<LOGL1, C, true>;
    static_assert(TL::NR >= 2, "direct column pass needs two rounds");
    __shared__ __align__(16) u32 lds[TL::LDS_WORDS];
    const int y = blockIdx.y;
    const ColSet &S = y < cnt0 ? s0 : s1;
    const int64_t a = y < cnt0 ? y : y - cnt0;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    const u32 *X = S.in + a * S.in_stride + c0;
    u32 *Y = S.out + a * S.out_stride + c0;
    const int64_t lim = S.in_len - c0;
    constexpr int64_t rows = TOPZERO ? TL::L / 2 : TL::L;
    if (lim >= (rows - 1) * n2 + C)      // every element round 0 reads is in range
        cm_col_dif_first<TL, TOPZERO, false>(lds, X, n2, lim, lv, md);
    else
        cm_col_dif_first<TL, TOPZERO, true>(lds, X, n2, lim, lv, md);
    __syncthreads();
    u32 *const t[1] = {lds};
    const MD mr = cm_rebind(md);
    cm_dif_rounds<TL, 1, TL::NR - 1, 1>(t, lv, mr);
    cm_col_dif_last<TL>(lds, Y, n2, lv, mr);
}

template <int LOGL1, int C, class MD>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T))
__attribute__((amdgpu_waves_per_eu(8, 8))) void k_col_inv_d(
    ColSet S, int logn2, const u32 *__restrict__ lvi, MD md) {
    using TL = Tile<LOGL1, C, true>;
    static_assert(TL::NR >= 2, "direct column pass needs two rounds");
    __shared__ __align__(16) u32 lds[TL::LDS_WORDS];
    const int64_t a = blockIdx.y;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    const u32 *X = S.in + a * S.in_stride + c0;
    u32 *Y = S.out + a * S.out_stride + c0;
    const int64_t lim = S.out_len - c0;
    cm_col_dit_first<TL>(lds, X, n2, lvi, md);
    __syncthreads();
    u32 *const t[1] = {lds};
    cm_dit_rounds<TL, TL::NR - 2, 1, 1>(t, lvi, md);
    if (lim >= (int64_t)(TL::L - 1) * n2 + C)
        cm_col_dit_last<TL, false>(lds, Y, n2, lim, lvi, md);
    else
        cm_col_dit_last<TL, true>(lds, Y, n2, lim, lvi, md);
}

// Column passes of the multiply for columns of at most 16 words (one
// round): each thread transforms one whole column in registers, straight
// from global memory and back, and a wave reads or writes 64 consecutive
// columns of a row (256-byte segments).  No LDS.  256 columns per workgroup.
constexpr int kColRegThreads = 256;

template <int LOGL1, bool TOPZERO, class MD>
__global__ __launch_bounds__(kColRegThreads) void k_col_fwd_r(
    ColSet s0, ColSet s1, int cnt0, int logn2, const u32 *__restrict__ lv, MD md) {
    static_assert(LOGL1 >= 1 && LOGL1 <= 4, "one round");
    constexpr int L = 1 << LOGL1, M = TOPZERO ? L / 2 : L;
    const int y = blockIdx.y;
    const ColSet &S = y < cnt0 ? s0 : s1;
    const int64_t a = y < cnt0 ? y : y - cnt0;
    const int64_t n2 = (int64_t)1 << logn2;
    const int64_t col = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kColRegThreads + threadIdx.x;
    const u32 *X = S.in + a * S.in_stride + col;
    u32 *Y = S.out + a * S.out_stride + col;
    const int64_t lim = S.in_len - col;
    u32 v[1][L];
#pragma unroll
    for (int m = 0; m < M; ++m) v[0][m] = (int64_t)m * n2 < lim ? X[(int64_t)m * n2] : 0u;
    cm_dif_stages<LOGL1, LOGL1, 0, 0, 1, TOPZERO>(v, 0, lv, md);
#pragma unroll
    for (int m = 0; m < L; ++m) Y[(int64_t)m * n2] = v[0][m];
}

template <int LOGL1, class MD>
__global__ __launch_bounds__(kColRegThreads) void k_col_inv_r(
    ColSet S, int logn2, const u32 *__restrict__ lvi, MD md) {
    static_assert(LOGL1 >= 1 && LOGL1 <= 4, "one round");
    constexpr int L = 1 << LOGL1;
    const int64_t a = blockIdx.y;
    const int64_t n2 = (int64_t)1 << logn2;
    const int64_t col = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kColRegThreads + threadIdx.x;
    const u32 *X = S.in + a * S.in_stride + col;
    u32 *Y = S.out + a * S.out_stride + col;
    const int64_t lim = S.out_len - col;
    u32 v[1][L];
#pragma unroll
    for (int m = 0; m < L; ++m) v[0][m] = X[(int64_t)m * n2];
    cm_dit_stages<LOGL1, LOGL1, 0, 0, 1>(v, 0, lvi, md);
#pragma unroll
    for (int m = 0; m < L; ++m)
        if ((int64_t)m * n2 < lim) Y[(int64_t)m * n2] = cm_canon(v[0][m], md);
}

// Row pass of the four-step multiply: row r of F and G (r = a * n1 + j, the
// row holding frequency k1 = brev(j)); elements i = t + T m are multiplied by
// the outer twiddle w_n^(k1 i) = rb[k1 T + t] * rc[16 k1 + m] on the way in,
// the two rows are multiplied in LDS (forward, pointwise, inverse), and the
// inverse outer twiddle is applied on the way out.  In place on F.
// Occupancy target: 2 workgroups per CU for 512-thread rows (4 waves/EU),
// 4 for 256-thread rows.
template <int LOGL2, class MD>
__global__ __launch_bounds__((Tile<LOGL2, 1, false>::T))
__attribute__((amdgpu_waves_per_eu(LOGL2 == 12 ? 4 : 4, 8))) void k_row_mul(
    u32 *F, const u32 *__restrict__ G, int logn1, const u32 *__restrict__ rb,
    const u32 *__restrict__ rc, const u32 *__restrict__ rbi, const u32 *__restrict__ rci,
    const u32 *__restrict__ lv, const u32 *__restrict__ lvi, MD md, u32 sc) {
    using TL = Tile<LOGL2, 1, false>;
    constexpr int T = TL::T;
    static_assert(TL::L == 16 * T, "row pass assumes 16 elements per thread");
    __shared__ u32 lds[2 * TL::LDS_WORDS];
    u32 *A = lds, *B = lds + TL::LDS_WORDS;
    const int64_t r = blockIdx.x;
    const int k1 = (int)cm_brev((u32)(r & ((1 << logn1) - 1)), logn1);
    u32 *Fr = F + r * TL::L;
    const u32 *Gr = G + r * TL::L;
    const int t = threadIdx.x;
    const u32 b = rb[(int64_t)k1 * T + t];
    const auto rf = cm_tile_rsrc(Fr), rg = cm_tile_rsrc(Gr);
    u32 fa[16], ga[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        fa[m] = cm_buf_ld(rf, (u32)t << 2, (u32)(T * m) << 2);
        ga[m] = cm_buf_ld(rg, (u32)t << 2, (u32)(T * m) << 2);
    }
    // one twiddle w_n^(k1 i) = rc b / 2^32 per element, shared by F and G; it
    // stays in [0, 2p), which keeps each product below p 2^32 (an operand
    // below 2p times a twiddle below 2p needs 4p < 2^32, p < 2^30, for the
    // lazy form; canonical operands need 2p < 2^32)
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const u32 tw = cm_mont_lazy(rc[k1 * 16 + m], b, md.p, md.np);
        fa[m] = cm_mulw(fa[m], tw, md);
        ga[m] = cm_mulw(ga[m], tw, md);
    }
    // elements t + T m are thread t's group of the first forward and the last
    // inverse round: those rounds run on the registers
    cm_row_multiply<TL>(A, B, fa, ga, lv, lvi, md, sc);
    const u32 bi = rbi[(int64_t)k1 * T + t];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const u32 c = rci[k1 * 16 + m];
        cm_buf_st(rf, (u32)t << 2, (u32)(T * m) << 2,
                  cm_mulw(cm_mont_lazy(fa[m], c, md.p, md.np), bi, md));
    }
}

// Row transforms for the natural-order stand-alone DFT: the workgroup takes
// the CR rows whose frequency index k1 = brev(row) is consecutive, and writes
// Y[k1 + k2 * n1] (out of place).  grid.y: arrays.
template <int LOGL2, int CR>
__global__ __launch_bounds__((Tile<LOGL2, CR, false>::T)) void k_row_out(
    const u32 *__restrict__ X0, u32 *__restrict__ Y0, int64_t stride, int logn1,
    const u32 *__restrict__ lv, Mod md, u32 sc) {
    using TL = Tile<LOGL2, CR, false>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const u32 *X = X0 + (int64_t)blockIdx.y * stride;
    u32 *Y = Y0 + (int64_t)blockIdx.y * stride;
    const int L = TL::L;
    const int64_t n1 = (int64_t)1 << logn1;
    const int k10 = blockIdx.x * CR;
    constexpr int PER = TL::WORDS / TL::T;
    u32 xa[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL2, i = idx & (L - 1);
        xa[r] = X[(int64_t)cm_brev(k10 + c, logn1) * L + i];
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * TL::T, c = idx >> LOGL2, i = idx & (L - 1);
        lds[TL::addr(c, i)] = xa[r];
    }
    __syncthreads();
    cm_tile_dif<TL>(lds, lv, md);
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int k2 = idx / CR, c = idx % CR;
        u32 v = lds[TL::addr(c, (int)cm_brev(k2, LOGL2))];
        if (sc) v = cm_mul_tw(v, sc, md.p, md.np);
        Y[k10 + c + k2 * n1] = v;
    }
}

// Natural-order column transforms DFT_m (x) I_ncols of a row-major
// 2^LOGM x ncols array, in place (the column half of the bivariate DFT,
// src/list_stockham.cu:1193-1201); output scaled by sc when sc != 0.
template <int LOGM, int C>
__global__ __launch_bounds__((Tile<LOGM, C, true>::T)) void k_col_ntt(
    u32 *__restrict__ X0, int64_t ncols, int64_t bstride, const u32 *__restrict__ lv, Mod md,
    u32 sc) {
    using TL = Tile<LOGM, C, true>;
    __shared__ u32 lds[TL::LDS_WORDS];
    u32 *X = X0 + (int64_t)blockIdx.y * bstride;   // grid.y: arrays bstride words apart
    const int64_t c0 = (int64_t)blockIdx.x * C;
    constexpr int PER = TL::WORDS / TL::T;
    u32 xa[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * TL::T, i = idx / C, c = idx % C;
        xa[r] = (c0 + c < ncols) ? X[i * ncols + c0 + c] : 0u;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) lds[TL::phys(threadIdx.x + r * TL::T)] = xa[r];
    __syncthreads();
    md = cm_rebind(md);
    cm_tile_dif<TL>(lds, lv, md);
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        if (c0 + c >= ncols) continue;
        u32 v = lds[TL::phys((int)cm_brev(i, LOGM) * C + c)];
        if (sc) v = cm_mul_tw(v, sc, md.p, md.np);
        X[i * ncols + c0 + c] = v;
    }
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_pointwise(u32 *__restrict__ Y, const u32 *__restrict__ X, int64_t n, u32 p,
                            u32 pinv, u32 r2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Y[i] = cm_mont(cm_mont(Y[i], X[i], p, pinv), r2, p, pinv);
}

__global__ void k_scale(u32 *__restrict__ X, int64_t n, u32 a, u32 aq, u32 p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        X[i] = cm_shoup(X[i], a, aq, p);
}

// W[j] = w^j: each thread raises w to its index by square-and-multiply on
// Montgomery residues (wm = w * 2^32 mod p, one = 2^32 mod p).
__global__ void k_powers(u32 *__restrict__ W, int64_t n, u32 wm, u32 one, u32 p, u32 pinv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        u32 r = one, x = wm;
        for (u64 e = (u64)i; e; e >>= 1) {
            if (e & 1) r = cm_mont(r, x, p, pinv);
            x = cm_mont(x, x, p, pinv);
        }
        W[i] = cm_mont(r, 1u, p, pinv);
    }
}

__global__ void k_fill(u32 *__restrict__ X, int64_t n, u32 v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        X[i] = v;
}

static int ew_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static Mod mod_of(const Family *fam) { return Mod{fam->p, fam->np}; }

// Montgomery form of the output scale `scale` (0 when scale == 1: no scaling)
static u32 scale_mont(u32 scale, u32 p) { return scale % p == 1u ? 0u : cm_host_mont(scale % p, p); }

// Pointwise constant of an N = 2^logn multiply: the Montgomery form of
// 2^32 / N mod p (see cm_tile_multiply).
u32 mul_scale(int logn, u32 p) {
    u32 ninv = cm_host_inv(cm_host_pow(2, logn, p), p);
    return cm_host_mont(cm_host_mul(cm_host_r1(p), ninv, p), p);
}

template <int LOGL>
static constexpr int rows_per_wg() { return LOGL >= 10 ? 1 : (1 << (LOGL >= 10 ? 0 : 10 - LOGL)); }

// Multiplies inside one workgroup keep their values in [0, 2p) between
// butterflies when p < 2^30 (ModLazy) and store canonical residues.
template <int LOGL, bool DESC>
static hipError_t mul_lds_t(const u32 *F, const u32 *G, u32 *H, const MulDesc *descs, int ndesc,
                            const MulDesc &one, const Family *fam, hipStream_t s) {
    constexpr int C = rows_per_wg<LOGL>();
    using TL = Tile<LOGL, C, false>;
    const int nb = (ndesc + C - 1) / C;
    ProfScope ps(kProfMulLds, s);
    if (fam->p < (1u << 30))
        hipLaunchKernelGGL((k_mul_lds<LOGL, C, DESC, ModLazy>), dim3(nb), dim3(TL::T), 0, s, F, G,
                           H, descs, ndesc, one, fam->lv, fam->lvi, ModLazy{fam->p, fam->np},
                           mul_scale(LOGL, fam->p));
    else
        hipLaunchKernelGGL((k_mul_lds<LOGL, C, DESC, Mod>), dim3(nb), dim3(TL::T), 0, s, F, G, H,
                           descs, ndesc, one, fam->lv, fam->lvi, mod_of(fam),
                           mul_scale(LOGL, fam->p));
    return hipGetLastError();
}

hipError_t launch_mul_lds(int logn, const u32 *F, const u32 *G, u32 *H, const MulDesc *d_descs,
                          int ndesc, const Family *fam, hipStream_t s) {
    MulDesc none{};
    return launch_mul_lds_one(logn, F, G, H, d_descs, ndesc, none, fam, s);
}

hipError_t launch_mul_lds_one(int logn, const u32 *F, const u32 *G, u32 *H,
                              const MulDesc *d_descs, int ndesc, const MulDesc &one,
                              const Family *fam, hipStream_t s) {
    if (ndesc <= 0) return hipSuccess;
    switch (logn) {
#define CM_CASE(K) case K: return mul_lds_t<K, true>(F, G, H, d_descs, ndesc, one, fam, s);
        CM_CASE(4) CM_CASE(5) CM_CASE(6) CM_CASE(7) CM_CASE(8) CM_CASE(9) CM_CASE(10)
        CM_CASE(11) CM_CASE(12) CM_CASE(13) CM_CASE(14)
#undef CM_CASE
    }
    return hipErrorInvalidValue;
}

template <int LOGL>
static hipError_t ntt_lds_t(u32 *X, int64_t m, const u32 *lv, u32 sc, const Family *fam,
                            hipStream_t s) {
    constexpr int C = rows_per_wg<LOGL>();
    using TL = Tile<LOGL, C, false>;
    const int64_t nb = (m + C - 1) / C;
    ProfScope ps(kProfNttLds, s);
    hipLaunchKernelGGL((k_ntt_lds<LOGL, C>), dim3((unsigned)nb), dim3(TL::T), 0, s, X, m, lv,
                       mod_of(fam), sc);
    return hipGetLastError();
}

hipError_t launch_ntt_lds(int k, u32 *X, int64_t m, bool inverse, u32 scale, const Family *fam,
                          hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const u32 *lv = inverse ? fam->lvi : fam->lv;
    const u32 sc = scale_mont(scale, fam->p);
    switch (k) {
#define CM_CASE(K) case K: return ntt_lds_t<K>(X, m, lv, sc, fam, s);
        CM_CASE(4) CM_CASE(5) CM_CASE(6) CM_CASE(7) CM_CASE(8) CM_CASE(9) CM_CASE(10)
        CM_CASE(11) CM_CASE(12) CM_CASE(13)
#undef CM_CASE
    }
    return hipErrorInvalidValue;
}

// column tile width: C columns x 2^LOGL1 rows <= WORDS words (at most 4096
// columns)
template <int LOGL1, int WORDS = 16384>
static constexpr int col_tile() {
    return (WORDS >> LOGL1) > 4096 ? 4096 : (WORDS >> LOGL1);
}

// 16-byte vector access to both arrays of a column set
static bool vec_ok(const ColSet &S) {
    return ((uintptr_t)S.in & 15) == 0 && ((uintptr_t)S.out & 15) == 0 && (S.in_stride & 3) == 0 &&
           (S.out_stride & 3) == 0;
}

// Words per column tile of the multiply's four-step column passes: 16384
// (two 1024-thread workgroups per CU) or 8192 (four 512-thread workgroups,
// half the row segment per tile); a tuned choice.
static int g_col_words = 16384;

bool set_col_tile_words(int words) {
    if (words != 8192 && words != 16384) return false;
    g_col_words = words;
    return true;
}

int col_tile_words() { return g_col_words; }

// Column passes of the multiply with direct global first/last rounds
// (k_col_fwd_d, k_col_inv_d) or with LDS staging (k_col_fwd, k_col_inv).
static bool g_col_direct = [] {
    const char *e = getenv("CUMODP_COL_DIRECT");
    return !(e && *e == '0');
}();

// MD = ModLazy (multiply only, values in [0, 2p) between the passes) runs
// the direct and register kernels only.
template <int LOGL1, int WORDS, class MD>
static hipError_t col_fwd_w(const ColSet &s0, const ColSet &s1, int cnt0, int cnt1, int logn2,
                            const u32 *lv, const u32 *olo, const u32 *ohi, MD md,
                            hipStream_t s) {
    constexpr int C = col_tile<LOGL1, WORDS>();
    using TL = Tile<LOGL1, C, true>;
    const int nb = (int)(((int64_t)1 << logn2) / C);
    ProfScope ps(kProfColFwd, s);
    const int64_t half = (int64_t)1 << (LOGL1 + logn2 - 1);
    const bool topzero = LOGL1 >= 2 && s0.in_len <= half && (cnt1 == 0 || s1.in_len <= half);
    if constexpr (!MD::LAZY) {
        if (olo) {
            hipLaunchKernelGGL((k_col_fwd<LOGL1, C, true, false>), dim3(nb, cnt0 + cnt1),
                               dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, olo, ohi, md);
            return hipGetLastError();
        }
    }
    if constexpr (LOGL1 <= 4) {
        if (g_col_direct && logn2 >= 8) {
            const dim3 g((unsigned)(((int64_t)1 << logn2) / kColRegThreads), cnt0 + cnt1);
            if (topzero)
                hipLaunchKernelGGL((k_col_fwd_r<LOGL1, true, MD>), g, dim3(kColRegThreads), 0, s,
                                   s0, s1, cnt0, logn2, lv, md);
            else
                hipLaunchKernelGGL((k_col_fwd_r<LOGL1, false, MD>), g, dim3(kColRegThreads), 0, s,
                                   s0, s1, cnt0, logn2, lv, md);
            return hipGetLastError();
        }
    }
    if constexpr (TL::NR >= 2) {
        if (g_col_direct) {
            if (topzero)
                hipLaunchKernelGGL((k_col_fwd_d<LOGL1, C, true, MD>), dim3(nb, cnt0 + cnt1),
                                   dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, md);
            else
                hipLaunchKernelGGL((k_col_fwd_d<LOGL1, C, false, MD>), dim3(nb, cnt0 + cnt1),
                                   dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, md);
            return hipGetLastError();
        }
    }
    if constexpr (MD::LAZY) {
        return hipErrorInvalidValue;
    } else if constexpr (C % 4 == 0) {
        if (vec_ok(s0) && (cnt1 == 0 || vec_ok(s1)) && topzero)
            hipLaunchKernelGGL((k_col_fwd<LOGL1, C, false, true, true>), dim3(nb, cnt0 + cnt1),
                               dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, olo, ohi, md);
        else if (vec_ok(s0) && (cnt1 == 0 || vec_ok(s1)))
            hipLaunchKernelGGL((k_col_fwd<LOGL1, C, false, true>), dim3(nb, cnt0 + cnt1),
                               dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, olo, ohi, md);
        else
            hipLaunchKernelGGL((k_col_fwd<LOGL1, C, false, false>), dim3(nb, cnt0 + cnt1),
                               dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, olo, ohi, md);
    } else {
        hipLaunchKernelGGL((k_col_fwd<LOGL1, C, false, false>), dim3(nb, cnt0 + cnt1),
                           dim3(TL::T), 0, s, s0, s1, cnt0, logn2, lv, olo, ohi, md);
    }
    return hipGetLastError();
}

template <int LOGL1, class MD>
static hipError_t col_fwd_t(const ColSet &s0, const ColSet &s1, int cnt0, int cnt1, int logn2,
                            const u32 *lv, const u32 *olo, const u32 *ohi, MD md,
                            hipStream_t s) {
    // the 8192-word tile serves the multiply only (stand-alone transforms
    // pass outer twiddles)
    if (!olo && g_col_words == 8192 && LOGL1 >= 9)
        return col_fwd_w<LOGL1, 8192>(s0, s1, cnt0, cnt1, logn2, lv, olo, ohi, md, s);
    return col_fwd_w<LOGL1, 16384>(s0, s1, cnt0, cnt1, logn2, lv, olo, ohi, md, s);
}

template <int LOGL1, int WORDS, class MD>
static hipError_t col_inv_w(const ColSet &S, int cnt, int logn2, const Family *fam, MD md,
                            hipStream_t s) {
    constexpr int C = col_tile<LOGL1, WORDS>();
    using TL = Tile<LOGL1, C, true>;
    const int nb = (int)(((int64_t)1 << logn2) / C);
    ProfScope ps(kProfColInv, s);
    if constexpr (LOGL1 <= 4) {
        if (g_col_direct && logn2 >= 8) {
            hipLaunchKernelGGL((k_col_inv_r<LOGL1, MD>),
                               dim3((unsigned)(((int64_t)1 << logn2) / kColRegThreads), cnt),
                               dim3(kColRegThreads), 0, s, S, logn2, fam->lvi, md);
            return hipGetLastError();
        }
    }
    if constexpr (TL::NR >= 2) {
        if (g_col_direct) {
            hipLaunchKernelGGL((k_col_inv_d<LOGL1, C, MD>), dim3(nb, cnt), dim3(TL::T), 0, s, S,
                               logn2, fam->lvi, md);
            return hipGetLastError();
        }
    }
    if constexpr (MD::LAZY) {
        return hipErrorInvalidValue;
    } else {
        if constexpr (C % 4 == 0) {
            if (vec_ok(S)) {
                hipLaunchKernelGGL((k_col_inv<LOGL1, C, true>), dim3(nb, cnt), dim3(TL::T), 0, s,
                                   S, logn2, fam->lvi, md);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_col_inv<LOGL1, C, false>), dim3(nb, cnt), dim3(TL::T), 0, s, S,
                           logn2, fam->lvi, md);
        return hipGetLastError();
    }
}

template <int LOGL1, class MD>
static hipError_t col_inv_t(const ColSet &S, int cnt, int logn2, const Family *fam, MD md,
                            hipStream_t s) {
    if (g_col_words == 8192 && LOGL1 >= 9)
        return col_inv_w<LOGL1, 8192>(S, cnt, logn2, fam, md, s);
    return col_inv_w<LOGL1, 16384>(S, cnt, logn2, fam, md, s);
}

template <int LOGL2>
static hipError_t row_out_t(const u32 *X, u32 *Y, int cnt, int64_t stride, int logn1,
                            const u32 *lv, u32 sc, const Family *fam, hipStream_t s) {
    constexpr int CR = LOGL2 >= 12 ? (1 << (14 - LOGL2)) : 4;   // 16384 words per tile
    using TL = Tile<LOGL2, CR, false>;
    const int nb = (int)(((int64_t)1 << logn1) / CR);
    ProfScope ps(kProfRowOut, s);
    hipLaunchKernelGGL((k_row_out<LOGL2, CR>), dim3(nb, cnt), dim3(TL::T), 0, s, X, Y, stride,
                       logn1, lv, mod_of(fam), sc);
    return hipGetLastError();
}

#define CM_K1_CASES(F, ...)                                                              \
    switch (k1) {                                                                        \
    case 1: return F<1>(__VA_ARGS__);   case 2: return F<2>(__VA_ARGS__);                \
    case 3: return F<3>(__VA_ARGS__);   case 4: return F<4>(__VA_ARGS__);                \
    case 5: return F<5>(__VA_ARGS__);   case 6: return F<6>(__VA_ARGS__);                \
    case 7: return F<7>(__VA_ARGS__);   case 8: return F<8>(__VA_ARGS__);                \
    case 9: return F<9>(__VA_ARGS__);   case 10: return F<10>(__VA_ARGS__);              \
    case 11: return F<11>(__VA_ARGS__); case 12: return F<12>(__VA_ARGS__);              \
    case 13: return F<13>(__VA_ARGS__);                                                  \
    }                                                                                    \
    return hipErrorInvalidValue;

template <class MD>
static hipError_t col_fwd(int k1, const ColSet &s0, const ColSet &s1, int cnt0, int cnt1,
                          int logn2, const u32 *lv, const u32 *olo, const u32 *ohi, MD md,
                          hipStream_t s) {
    CM_K1_CASES(col_fwd_t, s0, s1, cnt0, cnt1, logn2, lv, olo, ohi, md, s)
}

template <class MD>
static hipError_t col_inv(int k1, const ColSet &S, int cnt, int logn2, const Family *fam, MD md,
                          hipStream_t s) {
    CM_K1_CASES(col_inv_t, S, cnt, logn2, fam, md, s)
}

template <class MD>
static hipError_t mul_four_step_t(int k1, int k2, const FourStepIO &io, const Outer *o,
                                  const Family *fam, MD md, hipStream_t s) {
    const int64_t n = (int64_t)1 << (k1 + k2);
    ColSet sf{io.f_in, io.f_ws, io.in_stride, io.f_len, n, n};
    ColSet sg{io.g_in, io.g_ws, io.g_stride < 0 ? io.in_stride : io.g_stride, io.g_len, n, n};
    hipError_t e = col_fwd(k1, sf, sg, io.cnt, io.cnt, k2, fam->lv, nullptr, nullptr, md, s);
    if (e != hipSuccess) return e;
    // row pass: one row of F and G per workgroup (2^k1 * cnt rows); the
    // pointwise scale 2^32 / n is carried by the inverse row twiddles
    const int nrows = (int)(((int64_t)io.cnt << k1));
    const u32 sc = 0u;
    {
        ProfScope ps(kProfRowMul, s);
#define CM_ROW(K2)                                                                          \
    hipLaunchKernelGGL((k_row_mul<K2, MD>), dim3(nrows), dim3(Tile<K2, 1, false>::T), 0, s, \
                       io.f_ws, io.g_ws, k1, o->rb_f, o->rc_f, o->rb_i, o->rc_i, fam->lv,   \
                       fam->lvi, md, sc)
        if (k2 == 12) CM_ROW(12);
        else if (k2 == 13) CM_ROW(13);
        else CM_ROW(14);
#undef CM_ROW
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    ColSet sh{io.f_ws, io.h_out, n, n, io.out_stride, io.h_len};
    return col_inv(k1, sh, io.cnt, k2, fam, md, s);
}

// The multiply keeps its values in [0, 2p) between butterflies and passes
// (ModLazy) when p < 2^30 and the direct column kernels run; the last
// inverse round writes canonical residues.
hipError_t launch_mul_four_step(int k, const FourStepIO &io, const Family *fam, hipStream_t s) {
    if (k <= kMaxLdsLog || k > kMaxLog || k > fam->K) return hipErrorInvalidValue;
    int k1, k2;
    four_step_split(k, k1, k2);
    const Outer *o = outer_for(fam, k);
    if (!o) return hipErrorOutOfMemory;
    if (g_col_direct && fam->p < (1u << 30))
        return mul_four_step_t(k1, k2, io, o, fam, ModLazy{fam->p, fam->np}, s);
    return mul_four_step_t(k1, k2, io, o, fam, mod_of(fam), s);
}

hipError_t launch_mul_four_step_with(const FourStepIO &io, const Family *fam, const Outer *o,
                                     hipStream_t s) {
    if (!o || o->k > fam->K || o->k <= kMaxLdsLog) return hipErrorInvalidValue;
    if (g_col_direct && fam->p < (1u << 30))
        return mul_four_step_t(o->k1, o->k2, io, o, fam, ModLazy{fam->p, fam->np}, s);
    return mul_four_step_t(o->k1, o->k2, io, o, fam, mod_of(fam), s);
}

template <int LOGL>
static hipError_t rows_mul_t(u32 *F, u32 *G, int rows, const Family *fam, hipStream_t s) {
    const int L = 1 << LOGL;
    const MulDesc rows_desc{L, L, L, L, L, L, LOGL};
    return mul_lds_t<LOGL, false>(F, G, F, nullptr, rows, rows_desc, fam, s);
}

template <int LOGL>
static hipError_t mul_strided_t(const u32 *F, const u32 *G, u32 *H, const MulDesc &sd, int count,
                                const Family *fam, hipStream_t s) {
    return mul_lds_t<LOGL, false>(F, G, H, nullptr, count, sd, fam, s);
}

hipError_t launch_mul_strided(int logn, const u32 *F, int64_t fs, int lf, const u32 *G,
                              int64_t gs, int lg, u32 *H, int64_t hs, int lh, int count,
                              const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (logn < 4) logn = 4;
    // short operands, no wrap-around: the schoolbook kernel (same words)
    if (std::min(lf, lg) <= strided_plain_cutoff() && (int64_t)lf + lg - 1 <= ((int64_t)1 << logn) &&
        lh <= lf + lg - 1)
        return launch_plain_strided(F, fs, lf, G, gs, lg, H, hs, lh, count, fam->p, s);
    if (logn <= kMaxLdsLog) {
        const MulDesc sd{fs, gs, hs, lf, lg, lh, logn};
        switch (logn) {
#define CM_CASE(K) case K: return mul_strided_t<K>(F, G, H, sd, count, fam, s);
            CM_CASE(4) CM_CASE(5) CM_CASE(6) CM_CASE(7) CM_CASE(8) CM_CASE(9) CM_CASE(10)
            CM_CASE(11) CM_CASE(12) CM_CASE(13)
#undef CM_CASE
        }
        return hipErrorInvalidValue;
    }
    // four-step straight from the strided operands into the strided output,
    // in chunks whose two slabs stay within 1 GiB each
    const int64_t n = (int64_t)1 << logn;
    const int64_t chunk = std::max<int64_t>(1, ((int64_t)1 << 28) / n);
    for (int64_t c0 = 0; c0 < count; c0 += chunk) {
        const int cnt = (int)std::min<int64_t>(chunk, count - c0);
        u32 *sf = (u32 *)workspace(5, sizeof(u32) * (size_t)(n * cnt), s);
        u32 *sg = (u32 *)workspace(6, sizeof(u32) * (size_t)(n * cnt), s);
        if (!sf || !sg) return hipErrorOutOfMemory;
        FourStepIO io{F + c0 * fs, G + c0 * gs, lf, lg, fs, sf, sg, H + c0 * hs, lh, hs, cnt, gs};
        hipError_t e = launch_mul_four_step(logn, io, fam, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_rows_mul(int e, u32 *F, u32 *G, int rows, const Family *fam, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    switch (e) {
#define CM_CASE(K) case K: return rows_mul_t<K>(F, G, rows, fam, s);
        CM_CASE(4) CM_CASE(5) CM_CASE(6) CM_CASE(7) CM_CASE(8) CM_CASE(9) CM_CASE(10)
        CM_CASE(11) CM_CASE(12) CM_CASE(13)
#undef CM_CASE
    }
    const int64_t n = (int64_t)1 << e;
    FourStepIO io{F, G, n, n, n, F, G, F, n, n, rows};
    return launch_mul_four_step(e, io, fam, s);
}

hipError_t launch_ntt_four_step(int k, u32 *X, u32 *Y, int cnt, bool inverse, const Family *fam,
                                hipStream_t s) {
    if (k <= kMaxLdsLog || k > kMaxLog || k > fam->K) return hipErrorInvalidValue;
    int k1, k2;
    four_step_split(k, k1, k2);
    const Outer *o = outer_for(fam, k);
    if (!o) return hipErrorOutOfMemory;
    const int64_t n = (int64_t)1 << k;
    // the inverse transform is the forward transform with the inverse root
    const u32 *lv = inverse ? fam->lvi : fam->lv;
    const u32 *olo = inverse ? o->lo_i : o->lo_f, *ohi = inverse ? o->hi_i : o->hi_f;
    ColSet sx{X, X, n, n, n, n};
    hipError_t e = col_fwd(k1, sx, sx, cnt, 0, k2, lv, olo, ohi, mod_of(fam), s);
    if (e != hipSuccess) return e;
    const u32 sc =
        inverse ? scale_mont(cm_host_inv(cm_host_pow(2, k, fam->p), fam->p), fam->p) : 0u;
    if (k2 == 12) e = row_out_t<12>(X, Y, cnt, n, k1, lv, sc, fam, s);
    else if (k2 == 13) e = row_out_t<13>(X, Y, cnt, n, k1, lv, sc, fam, s);
    else e = row_out_t<14>(X, Y, cnt, n, k1, lv, sc, fam, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(X, Y, sizeof(u32) * (size_t)(n * cnt), hipMemcpyDeviceToDevice, s);
}

template <int LOGM>
static hipError_t col_ntt_t(u32 *X, int64_t ncols, int nblk, int64_t bstride, u32 sc,
                            const Family *fam, hipStream_t s) {
    constexpr int C = col_tile<LOGM>();
    using TL = Tile<LOGM, C, true>;
    const int64_t nb = (ncols + C - 1) / C;
    ProfScope ps(kProfColNtt, s);
    hipLaunchKernelGGL((k_col_ntt<LOGM, C>), dim3((unsigned)nb, (unsigned)nblk), dim3(TL::T), 0,
                       s, X, ncols, bstride, fam->lv, mod_of(fam), sc);
    return hipGetLastError();
}

hipError_t launch_col_ntt(int em, u32 *X, int64_t ncols, u32 scale, const Family *fam,
                          hipStream_t s, int nblk, int64_t bstride) {
    if (nblk <= 0 || nblk > 65535) return hipErrorInvalidValue;
    const int k1 = em;
    const u32 sc = scale_mont(scale, fam->p);
    CM_K1_CASES(col_ntt_t, X, ncols, nblk, bstride, sc, fam, s)
}

hipError_t launch_pointwise(u32 *Y, const u32 *X, int64_t n, u32 p, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    u32 r1 = cm_host_r1(p);
    u32 r2 = cm_host_mul(r1, r1, p);
    hipLaunchKernelGGL(k_pointwise, dim3(ew_blocks(n)), dim3(256), 0, s, Y, X, n, p,
                       cm_host_pinv(p), r2);
    return hipGetLastError();
}

hipError_t launch_scale(u32 *X, int64_t n, u32 a, u32 p, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    a %= p;
    hipLaunchKernelGGL(k_scale, dim3(ew_blocks(n)), dim3(256), 0, s, X, n, a,
                       cm_host_shoup_q(a, p), p);
    return hipGetLastError();
}

hipError_t launch_powers(u32 *W, int64_t n, u32 w, u32 p, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    u32 one = cm_host_r1(p);
    hipLaunchKernelGGL(k_powers, dim3(ew_blocks(n)), dim3(256), 0, s, W, n,
                       cm_host_mul(w % p, one, p), one, p, cm_host_pinv(p));
    return hipGetLastError();
}

hipError_t launch_fill(u32 *X, int64_t n, u32 v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(ew_blocks(n)), dim3(256), 0, s, X, n, v);
    return hipGetLastError();
}

}  // namespace cm
