// kernels_ntt.hip -- NTT, NTT-multiply and elementwise kernels for gfx950.
//
// Replaces the reference's one-stage-per-launch radix-2 Stockham passes
// (src/stockham.cu:75-397, src/list_stockham.cu:108-576, src/fft_aux.cu:70-322)
// with:
//   * k_mul_lds   -- a whole multiply (forward F, forward G, pointwise,
//                    inverse, scaling) inside one workgroup for 2^k <= 2^13;
//   * k_col_fwd / k_row_mul / k_col_inv -- a three-kernel four-step multiply
//                    for larger sizes: column transforms with the outer
//                    twiddle fused, fused row transforms + pointwise product +
//                    inverse row transforms, inverse column transforms;
//   * k_ntt_lds / k_row_out -- natural-order stand-alone transforms.
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
template <int LOGL, int C, bool DESC>
__global__ __launch_bounds__((Tile<LOGL, C, false>::T)) void k_mul_lds(
    const u32 *F, const u32 *__restrict__ G, u32 *H,  // H may alias F
    const MulDesc *__restrict__ descs, int ndesc, MulDesc one, const uint2 *__restrict__ tw,
    const uint2 *__restrict__ twi, int tws, u32 p, u32 pinv, uint2 sc) {
    // DESC: descriptors from `descs` (or the by-value `one` when descs is
    // null); otherwise row d of F, G, H at offset d * L, full length.
    using TL = Tile<LOGL, C, false>;
    using LR = LastRound<TL>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const int L = TL::L;

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
            off_f[c] = off_g[c] = off_h[c] = (int64_t)d * L;
            lf[c] = lg[c] = lh[c] = d < ndesc ? L : 0;
        }
    }

#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int i = threadIdx.x; i < L; i += TL::T)
            lds[TL::addr(c, i)] = i < lf[c] ? F[off_f[c] + i] : 0u;
    __syncthreads();
    cm_dif_rounds<TL, 0, TL::NR - 1>(lds, tw, tws, p);
    u32 fr[LR::GPT][1 << LR::R];
    LR::dif_keep(lds, fr, tw, tws, p);
    __syncthreads();

#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int i = threadIdx.x; i < L; i += TL::T)
            lds[TL::addr(c, i)] = i < lg[c] ? G[off_g[c] + i] : 0u;
    __syncthreads();
    cm_dif_rounds<TL, 0, TL::NR - 1>(lds, tw, tws, p);
    LR::dif_mul_dit(lds, fr, tw, twi, tws, p, pinv, sc);
    cm_dit_rounds<TL, TL::NR - 2, 0>(lds, twi, tws, p);

#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int i = threadIdx.x; i < lh[c]; i += TL::T)
            H[off_h[c] + i] = i < L ? lds[TL::addr(c, i)] : 0u;
}

// ---------------------------------------------------------------------------
// natural-order forward DFT of rows inside one workgroup
// ---------------------------------------------------------------------------
template <int LOGL, int C>
__global__ __launch_bounds__((Tile<LOGL, C, false>::T)) void k_ntt_lds(
    u32 *__restrict__ X, int64_t m, const uint2 *__restrict__ tw, int tws, u32 p, uint2 sc) {
    using TL = Tile<LOGL, C, false>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const int L = TL::L;
    const int64_t r0 = (int64_t)blockIdx.x * C;
#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int i = threadIdx.x; i < L; i += TL::T)
            lds[TL::addr(c, i)] = (r0 + c < m) ? X[(r0 + c) * L + i] : 0u;
    __syncthreads();
    cm_tile_dif<TL>(lds, tw, tws, p);
    const bool scale = sc.x != 1u;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (r0 + c >= m) break;
        for (int i = threadIdx.x; i < L; i += TL::T) {
            u32 v = lds[TL::addr(c, (int)cm_brev(i, LOGL))];
            if (scale) v = cm_shoup(v, sc.x, sc.y, p);
            X[(r0 + c) * L + i] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// four-step pieces: X viewed as 2^k1 rows x 2^k2 columns
// ---------------------------------------------------------------------------
// Forward column transforms (DIF, bit-reversed rows) followed by the outer
// twiddle w_n^(i2 * k1); in place.  grid.x: column tiles, grid.y: arrays
// (the first cnt0 at X0 + y*stride, the rest at X1 + (y-cnt0)*stride).
template <int LOGL1, int C>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T)) void k_col_fwd(
    u32 *__restrict__ X0, int cnt0, u32 *__restrict__ X1, int64_t stride, int logn2,
    const uint2 *__restrict__ tw, int tws, const uint2 *__restrict__ olo,
    const uint2 *__restrict__ ohi, u32 p) {
    using TL = Tile<LOGL1, C, true>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const int y = blockIdx.y;
    u32 *X = y < cnt0 ? X0 + (int64_t)y * stride : X1 + (int64_t)(y - cnt0) * stride;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        lds[cm_pad(idx)] = X[i * n2 + c0 + c];
    }
    __syncthreads();
    cm_tile_dif<TL>(lds, tw, tws, p);
    constexpr int LOM = (1 << kOuterLoBits) - 1;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        const u32 e = (u32)(c0 + c) * cm_brev(i, LOGL1);
        const uint2 a = olo[e & LOM], b = ohi[e >> kOuterLoBits];
        u32 v = cm_shoup_lazy(lds[cm_pad(idx)], a.x, a.y, p);
        v = cm_shoup(v, b.x, b.y, p);
        X[i * n2 + c0 + c] = v;
    }
}

// Inverse: outer twiddle w_n^-(i2 * k1), then inverse column transforms
// (DIT, bit-reversed rows in, natural rows out); in place.
template <int LOGL1, int C>
__global__ __launch_bounds__((Tile<LOGL1, C, true>::T)) void k_col_inv(
    u32 *__restrict__ X0, int64_t stride, int logn2, const uint2 *__restrict__ twi, int tws,
    const uint2 *__restrict__ oilo, const uint2 *__restrict__ oihi, u32 p) {
    using TL = Tile<LOGL1, C, true>;
    __shared__ u32 lds[TL::LDS_WORDS];
    u32 *X = X0 + (int64_t)blockIdx.y * stride;
    const int64_t n2 = (int64_t)1 << logn2;
    const int c0 = xcd_remap(blockIdx.x, gridDim.x) * C;
    constexpr int LOM = (1 << kOuterLoBits) - 1;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        const u32 e = (u32)(c0 + c) * cm_brev(i, LOGL1);
        const uint2 a = oilo[e & LOM], b = oihi[e >> kOuterLoBits];
        u32 v = cm_shoup_lazy(X[i * n2 + c0 + c], a.x, a.y, p);
        lds[cm_pad(idx)] = cm_shoup(v, b.x, b.y, p);
    }
    __syncthreads();
    cm_tile_dit<TL>(lds, twi, tws, p);
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        X[i * n2 + c0 + c] = lds[cm_pad(idx)];
    }
}

// Row transforms for the natural-order stand-alone DFT: the workgroup takes
// the CR rows whose frequency index k1 = brev(row) is consecutive, and writes
// Y[k1 + k2 * n1] (out of place).  grid.y: arrays.
template <int LOGL2, int CR>
__global__ __launch_bounds__((Tile<LOGL2, CR, false>::T)) void k_row_out(
    const u32 *__restrict__ X0, u32 *__restrict__ Y0, int64_t stride, int logn1,
    const uint2 *__restrict__ tw, int tws, u32 p, uint2 sc) {
    using TL = Tile<LOGL2, CR, false>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const u32 *X = X0 + (int64_t)blockIdx.y * stride;
    u32 *Y = Y0 + (int64_t)blockIdx.y * stride;
    const int L = TL::L;
    const int64_t n1 = (int64_t)1 << logn1;
    const int k10 = blockIdx.x * CR;
#pragma unroll
    for (int c = 0; c < CR; ++c) {
        const int64_t row = cm_brev(k10 + c, logn1);
        for (int i = threadIdx.x; i < L; i += TL::T) lds[TL::addr(c, i)] = X[row * L + i];
    }
    __syncthreads();
    cm_tile_dif<TL>(lds, tw, tws, p);
    const bool scale = sc.x != 1u;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int k2 = idx / CR, c = idx % CR;
        u32 v = lds[TL::addr(c, (int)cm_brev(k2, LOGL2))];
        if (scale) v = cm_shoup(v, sc.x, sc.y, p);
        Y[k10 + c + k2 * n1] = v;
    }
}

// Natural-order column transforms DFT_m (x) I_ncols of a row-major
// 2^LOGM x ncols array, in place (the column half of the bivariate DFT,
// src/list_stockham.cu:1193-1201); output scaled by sc when sc.x != 1.
template <int LOGM, int C>
__global__ __launch_bounds__((Tile<LOGM, C, true>::T)) void k_col_ntt(
    u32 *__restrict__ X, int64_t ncols, const uint2 *__restrict__ tw, int tws, u32 p, uint2 sc) {
    using TL = Tile<LOGM, C, true>;
    __shared__ u32 lds[TL::LDS_WORDS];
    const int64_t c0 = (int64_t)blockIdx.x * C;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        lds[cm_pad(idx)] = (c0 + c < ncols) ? X[i * ncols + c0 + c] : 0u;
    }
    __syncthreads();
    cm_tile_dif<TL>(lds, tw, tws, p);
    const bool scale = sc.x != 1u;
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += TL::T) {
        const int i = idx / C, c = idx % C;
        if (c0 + c >= ncols) continue;
        u32 v = lds[cm_pad((int)cm_brev(i, LOGM) * C + c)];
        if (scale) v = cm_shoup(v, sc.x, sc.y, p);
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
static uint2 shoup_pair(u32 w, u32 p) { return make_uint2(w, cm_host_shoup_q(w, p)); }

// Shoup constant folding Montgomery's 2^-32 and 1/N: 2^32 / N mod p.
static uint2 mul_scale(int logn, u32 p) {
    u32 ninv = cm_host_inv(cm_host_pow(2, logn, p), p);
    return shoup_pair(cm_host_mul(cm_host_r1(p), ninv, p), p);
}

template <int LOGL>
static constexpr int rows_per_wg() { return LOGL >= 10 ? 1 : (1 << (LOGL >= 10 ? 0 : 10 - LOGL)); }

template <int LOGL, bool DESC>
static hipError_t mul_lds_t(const u32 *F, const u32 *G, u32 *H, const MulDesc *descs, int ndesc,
                            const MulDesc &one, const Family *fam, hipStream_t s) {
    constexpr int C = rows_per_wg<LOGL>();
    using TL = Tile<LOGL, C, false>;
    const int tws = fam->LT - LOGL;
    const int nb = (ndesc + C - 1) / C;
    ProfScope ps(kProfMulLds, s);
    hipLaunchKernelGGL((k_mul_lds<LOGL, C, DESC>), dim3(nb), dim3(TL::T), 0, s, F, G, H, descs,
                       ndesc, one, fam->fwd, fam->inv, tws, fam->p, fam->pinv,
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
        CM_CASE(11) CM_CASE(12) CM_CASE(13)
#undef CM_CASE
    }
    return hipErrorInvalidValue;
}

template <int LOGL>
static hipError_t ntt_lds_t(u32 *X, int64_t m, const uint2 *tw, uint2 sc, const Family *fam,
                            hipStream_t s) {
    constexpr int C = rows_per_wg<LOGL>();
    using TL = Tile<LOGL, C, false>;
    const int64_t nb = (m + C - 1) / C;
    ProfScope ps(kProfNttLds, s);
    hipLaunchKernelGGL((k_ntt_lds<LOGL, C>), dim3((unsigned)nb), dim3(TL::T), 0, s, X, m, tw,
                       fam->LT - LOGL, fam->p, sc);
    return hipGetLastError();
}

hipError_t launch_ntt_lds(int k, u32 *X, int64_t m, const uint2 *tw, uint2 sc, const Family *fam,
                          hipStream_t s) {
    if (m <= 0) return hipSuccess;
    switch (k) {
#define CM_CASE(K) case K: return ntt_lds_t<K>(X, m, tw, sc, fam, s);
        CM_CASE(4) CM_CASE(5) CM_CASE(6) CM_CASE(7) CM_CASE(8) CM_CASE(9) CM_CASE(10)
        CM_CASE(11) CM_CASE(12) CM_CASE(13)
#undef CM_CASE
    }
    return hipErrorInvalidValue;
}

// column tile width: C columns x 2^LOGL1 rows <= 16384 words
template <int LOGL1>
static constexpr int col_tile() {
    return (16384 >> LOGL1) > 4096 ? 4096 : (16384 >> LOGL1);
}

template <int LOGL1>
static hipError_t col_fwd_t(u32 *X0, int cnt0, u32 *X1, int cnt1, int64_t stride, int logn2,
                            const Family *fam, const Outer *o, hipStream_t s) {
    constexpr int C = col_tile<LOGL1>();
    using TL = Tile<LOGL1, C, true>;
    const int nb = (int)(((int64_t)1 << logn2) / C);
    ProfScope ps(kProfColFwd, s);
    hipLaunchKernelGGL((k_col_fwd<LOGL1, C>), dim3(nb, cnt0 + cnt1), dim3(TL::T), 0, s, X0,
                       cnt0, X1, stride, logn2, fam->fwd, fam->LT - LOGL1, o->lo_f, o->hi_f,
                       fam->p);
    return hipGetLastError();
}

template <int LOGL1>
static hipError_t col_inv_t(u32 *X, int cnt, int64_t stride, int logn2, const Family *fam,
                            const Outer *o, hipStream_t s) {
    constexpr int C = col_tile<LOGL1>();
    using TL = Tile<LOGL1, C, true>;
    const int nb = (int)(((int64_t)1 << logn2) / C);
    ProfScope ps(kProfColInv, s);
    hipLaunchKernelGGL((k_col_inv<LOGL1, C>), dim3(nb, cnt), dim3(TL::T), 0, s, X, stride,
                       logn2, fam->inv, fam->LT - LOGL1, o->lo_i, o->hi_i, fam->p);
    return hipGetLastError();
}

template <int LOGL2>
static hipError_t row_out_t(const u32 *X, u32 *Y, int cnt, int64_t stride, int logn1,
                            const uint2 *tw, uint2 sc, const Family *fam, hipStream_t s) {
    constexpr int CR = LOGL2 >= 12 ? (1 << (14 - LOGL2)) : 4;
    using TL = Tile<LOGL2, CR, false>;
    const int nb = (int)(((int64_t)1 << logn1) / CR);
    ProfScope ps(kProfRowOut, s);
    hipLaunchKernelGGL((k_row_out<LOGL2, CR>), dim3(nb, cnt), dim3(TL::T), 0, s, X, Y, stride,
                       logn1, tw, fam->LT - LOGL2, fam->p, sc);
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

static hipError_t col_fwd(int k1, u32 *X0, int cnt0, u32 *X1, int cnt1, int64_t stride,
                          int logn2, const Family *fam, const Outer *o, hipStream_t s) {
    CM_K1_CASES(col_fwd_t, X0, cnt0, X1, cnt1, stride, logn2, fam, o, s)
}

static hipError_t col_inv(int k1, u32 *X, int cnt, int64_t stride, int logn2, const Family *fam,
                          const Outer *o, hipStream_t s) {
    CM_K1_CASES(col_inv_t, X, cnt, stride, logn2, fam, o, s)
}

hipError_t launch_mul_four_step(int k, u32 *F, u32 *G, int cnt, const Family *fam,
                                hipStream_t s) {
    if (k <= kMaxLdsLog || k > kMaxLog || k > fam->K) return hipErrorInvalidValue;
    int k1, k2;
    four_step_split(k, k1, k2);
    const Outer *o = outer_for(fam, k);
    if (!o) return hipErrorOutOfMemory;
    const int64_t n = (int64_t)1 << k;
    hipError_t e = col_fwd(k1, F, cnt, G, cnt, n, k2, fam, o, s);
    if (e != hipSuccess) return e;
    // row multiply: one row of F and G per workgroup (2^k1 * cnt rows)
    const int nrows = (int)(((int64_t)cnt << k1));
    uint2 sc = mul_scale(k, fam->p);
    const int tws = fam->LT - k2;
    {
    ProfScope ps(kProfRowMul, s);
    if (k2 == 12) {
        using TL = Tile<12, 1, false>;
        hipLaunchKernelGGL((k_mul_lds<12, 1, false>), dim3(nrows), dim3(TL::T), 0, s, F, G, F,
                           nullptr, nrows, MulDesc{}, fam->fwd, fam->inv, tws, fam->p,
                           fam->pinv, sc);
    } else {
        using TL = Tile<13, 1, false>;
        hipLaunchKernelGGL((k_mul_lds<13, 1, false>), dim3(nrows), dim3(TL::T), 0, s, F, G, F,
                           nullptr, nrows, MulDesc{}, fam->fwd, fam->inv, tws, fam->p,
                           fam->pinv, sc);
    }
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return col_inv(k1, F, cnt, n, k2, fam, o, s);
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
    Outer oi = *o;
    const uint2 *tw = fam->fwd;
    if (inverse) {
        oi.lo_f = o->lo_i; oi.hi_f = o->hi_i;
        tw = fam->inv;
    }
    Family fi = *fam;
    fi.fwd = const_cast<uint2 *>(tw);
    hipError_t e = col_fwd(k1, X, cnt, X, 0, n, k2, &fi, &oi, s);
    if (e != hipSuccess) return e;
    uint2 sc = make_uint2(1u, 0u);
    if (inverse) sc = shoup_pair(cm_host_inv(cm_host_pow(2, k, fam->p), fam->p), fam->p);
    if (k2 == 12) e = row_out_t<12>(X, Y, cnt, n, k1, tw, sc, fam, s);
    else e = row_out_t<13>(X, Y, cnt, n, k1, tw, sc, fam, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(X, Y, sizeof(u32) * (size_t)(n * cnt), hipMemcpyDeviceToDevice, s);
}

template <int LOGM>
static hipError_t col_ntt_t(u32 *X, int64_t ncols, const uint2 *tw, uint2 sc, const Family *fam,
                            hipStream_t s) {
    constexpr int C = col_tile<LOGM>();
    using TL = Tile<LOGM, C, true>;
    const int64_t nb = (ncols + C - 1) / C;
    ProfScope ps(kProfColNtt, s);
    hipLaunchKernelGGL((k_col_ntt<LOGM, C>), dim3((unsigned)nb), dim3(TL::T), 0, s, X, ncols, tw,
                       fam->LT - LOGM, fam->p, sc);
    return hipGetLastError();
}

hipError_t launch_col_ntt(int em, u32 *X, int64_t ncols, const uint2 *tw, uint2 sc,
                          const Family *fam, hipStream_t s) {
    const int k1 = em;
    CM_K1_CASES(col_ntt_t, X, ncols, tw, sc, fam, s)
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
