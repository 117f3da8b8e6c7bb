// ntt_core.h -- in-LDS number-theoretic transform building blocks for gfx950.
//
// A tile holds C independent length-L vectors (L = 2^LOGL) in LDS, either
// row-major (C rows of L contiguous words) or column-major (L rows of C
// words, i.e. C columns of a larger matrix).  Transforms run in "rounds":
// one round applies R consecutive radix-2 stages to groups of 2^R elements
// that each thread holds in registers, so a 4096-point transform is three
// register-resident radix-16 passes with two LDS exchanges in between.
//
// Forward transforms are decimation in frequency (natural order in,
// bit-reversed order out); inverse transforms are the exact stage-by-stage
// inverse (decimation in time, bit-reversed in, natural out) with inverse
// twiddles.  The multiply path therefore never permutes data: the pointwise
// product is taken in bit-reversed order.  Stand-alone transforms that must
// return the natural-order DFT Y_j = sum_i X_i w^(ij) (the output of the
// reference's Stockham kernels, src/stockham.cu:40-46 and :413-453) read
// the tile back through a bit reversal.
//
// DIF stage s (0-based) pairs positions q and q + d, d = L >> (s+1), inside
// blocks of 2d words, and maps (x, y) -> (x + y, (x - y) * w_L^((q mod d) 2^s)).
//
// Twiddle tables hold uint2 {w^j, floor(w^j 2^32 / p)} for a root of order
// Ltab >= L; entry j << tws of the table is w_L^j with tws = log2(Ltab/L).
#pragma once

#include "modarith.h"

#define CM_DEV __device__ __forceinline__

// One pad word after every 16 LDS words keeps the stride-16 access of the
// last round (and the column tiles) free of bank conflicts.
CM_DEV int cm_pad(int i) { return i + (i >> 4); }

CM_DEV u32 cm_brev(u32 x, int bits) { return __brev(x) >> (32 - bits); }

// (x, y) -> (x + y, (x - y) w): the difference enters Shoup unreduced.
CM_DEV void cm_dif_bfly(u32 &x, u32 &y, uint2 w, u32 p) {
    u32 s = cm_add(x, y, p);
    u32 d = x - y + p;                 // (0, 2p)
    y = cm_shoup(d, w.x, w.y, p);
    x = s;
}

CM_DEV void cm_dif_bfly1(u32 &x, u32 &y, u32 p) {
    u32 s = cm_add(x, y, p);
    y = cm_sub(x, y, p);
    x = s;
}

// (a, b) -> (a + b w, a - b w)
CM_DEV void cm_dit_bfly(u32 &a, u32 &b, uint2 w, u32 p) {
    u32 t = cm_shoup(b, w.x, w.y, p);
    b = cm_sub(a, t, p);
    a = cm_add(a, t, p);
}

CM_DEV void cm_dit_bfly1(u32 &a, u32 &b, u32 p) {
    u32 t = b;
    b = cm_sub(a, t, p);
    a = cm_add(a, t, p);
}

template <int LOGL_, int C_, bool COLS_, int T_ = (C_ << LOGL_) / 16>
struct Tile {
    static constexpr int LOGL = LOGL_;
    static constexpr int L = 1 << LOGL_;
    static constexpr int C = C_;
    static constexpr bool COLS = COLS_;
    static constexpr int WORDS = C_ * L;
    static constexpr int LDS_WORDS = WORDS + (WORDS >> 4);
    static constexpr int T = T_;                   // threads per workgroup
    static constexpr int NR = (LOGL_ + 3) / 4;     // rounds
    static constexpr int rsz(int j) { return LOGL_ / NR + (j < LOGL_ % NR ? 1 : 0); }
    static constexpr int s0(int j) {
        int s = 0;
        for (int i = 0; i < j; ++i) s += rsz(i);
        return s;
    }
    // LDS address of element i of vector c
    static CM_DEV int addr(int c, int i) { return cm_pad(COLS_ ? i * C_ + c : c * L + i); }
    // group g of a round with 2^(LOGL-R) groups per vector -> (vector, index)
    template <int R>
    static CM_DEV void split(int g, int &c, int &gi) {
        if (COLS_) { c = g % C_; gi = g / C_; }
        else { c = g >> (LOGL_ - R); gi = g & ((1 << (LOGL_ - R)) - 1); }
    }
};

// R DIF stages on a register group (stages S0 .. S0+R-1).
template <int R, int S0, int DLOG>
CM_DEV void cm_dif_stages(u32 (&v)[1 << R], int gi_lo, const uint2 *__restrict__ tw, int tws,
                          u32 p) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int h = 1 << (R - 1 - t);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) {
            if (m & h) continue;
            const int mm = (m & (h - 1)) << DLOG;
            if (DLOG == 0 && mm == 0) {
                cm_dif_bfly1(v[m], v[m + h], p);
            } else {
                const int e = ((DLOG == 0 ? 0 : gi_lo) + mm) << (S0 + t);
                cm_dif_bfly(v[m], v[m + h], tw[e << tws], p);
            }
        }
    }
}

// Inverse of cm_dif_stages with the inverse twiddle table.
template <int R, int S0, int DLOG>
CM_DEV void cm_dit_stages(u32 (&v)[1 << R], int gi_lo, const uint2 *__restrict__ twi, int tws,
                          u32 p) {
#pragma unroll
    for (int t = R - 1; t >= 0; --t) {
        const int h = 1 << (R - 1 - t);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) {
            if (m & h) continue;
            const int mm = (m & (h - 1)) << DLOG;
            if (DLOG == 0 && mm == 0) {
                cm_dit_bfly1(v[m], v[m + h], p);
            } else {
                const int e = ((DLOG == 0 ? 0 : gi_lo) + mm) << (S0 + t);
                cm_dit_bfly(v[m], v[m + h], twi[e << tws], p);
            }
        }
    }
}

template <class TL, int J>
struct Round {
    static constexpr int R = TL::rsz(J);
    static constexpr int S0 = TL::s0(J);
    static constexpr int DLOG = TL::LOGL - S0 - R;
    static constexpr int NG = TL::WORDS >> R;   // groups in the tile
    static constexpr int GPT = NG / TL::T;      // groups per thread
    static_assert(GPT >= 1 && GPT * TL::T == NG, "round does not tile the workgroup");

    struct G {
        int c, gi_lo, i0;
        CM_DEV explicit G(int g) {
            int gi;
            TL::template split<R>(g, c, gi);
            gi_lo = gi & ((1 << DLOG) - 1);
            i0 = ((gi >> DLOG) << (DLOG + R)) | gi_lo;
        }
        CM_DEV int a(int m) const { return TL::addr(c, i0 + (m << DLOG)); }
    };

    static CM_DEV void load(const u32 *lds, const G &g, u32 (&v)[1 << R]) {
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) v[m] = lds[g.a(m)];
    }
    static CM_DEV void store(u32 *lds, const G &g, const u32 (&v)[1 << R]) {
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) lds[g.a(m)] = v[m];
    }
    // full forward round, LDS -> LDS
    static CM_DEV void dif(u32 *lds, const uint2 *tw, int tws, u32 p) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            G g(threadIdx.x + q * TL::T);
            u32 v[1 << R];
            load(lds, g, v);
            cm_dif_stages<R, S0, DLOG>(v, g.gi_lo, tw, tws, p);
            store(lds, g, v);
        }
    }
    // full inverse round, LDS -> LDS
    static CM_DEV void dit(u32 *lds, const uint2 *twi, int tws, u32 p) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            G g(threadIdx.x + q * TL::T);
            u32 v[1 << R];
            load(lds, g, v);
            cm_dit_stages<R, S0, DLOG>(v, g.gi_lo, twi, tws, p);
            store(lds, g, v);
        }
    }
};

// Forward rounds [J, J1) with a barrier after each.
template <class TL, int J, int J1>
CM_DEV void cm_dif_rounds(u32 *lds, const uint2 *tw, int tws, u32 p) {
    if constexpr (J < J1) {
        Round<TL, J>::dif(lds, tw, tws, p);
        __syncthreads();
        cm_dif_rounds<TL, J + 1, J1>(lds, tw, tws, p);
    }
}

// Inverse rounds J, J-1, ..., J0 (J >= J0) with a barrier after each.
template <class TL, int J, int J0>
CM_DEV void cm_dit_rounds(u32 *lds, const uint2 *twi, int tws, u32 p) {
    if constexpr (J >= J0) {
        Round<TL, J>::dit(lds, twi, tws, p);
        __syncthreads();
        cm_dit_rounds<TL, J - 1, J0>(lds, twi, tws, p);
    }
}

// Full forward transform of every vector in the tile (leaves bit-reversed
// order, barrier included).
template <class TL>
CM_DEV void cm_tile_dif(u32 *lds, const uint2 *tw, int tws, u32 p) {
    cm_dif_rounds<TL, 0, TL::NR>(lds, tw, tws, p);
}

template <class TL>
CM_DEV void cm_tile_dit(u32 *lds, const uint2 *twi, int tws, u32 p) {
    cm_dit_rounds<TL, TL::NR - 1, 0>(lds, twi, tws, p);
}

// Pointwise product of two transformed tiles held as the last forward round
// (fr: first operand kept in registers, lds: second operand), scaled by the
// Shoup constant sc (= 2^32 / N mod p, cancelling Montgomery's 2^-32 and
// the inverse transform's factor N), followed by the first inverse round;
// the result is written back to LDS.  Barrier included.
template <class TL>
struct LastRound {
    using RD = Round<TL, TL::NR - 1>;
    static constexpr int R = RD::R;
    static constexpr int GPT = RD::GPT;
    // forward last round of the first operand, kept in registers
    static CM_DEV void dif_keep(const u32 *lds, u32 (&fr)[GPT][1 << R], const uint2 *tw, int tws,
                                u32 p) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            typename RD::G g(threadIdx.x + q * TL::T);
            RD::load(lds, g, fr[q]);
            cm_dif_stages<R, RD::S0, RD::DLOG>(fr[q], g.gi_lo, tw, tws, p);
        }
    }
    static CM_DEV void dif_mul_dit(u32 *lds, const u32 (&fr)[GPT][1 << R], const uint2 *tw,
                                   const uint2 *twi, int tws, u32 p, u32 pinv, uint2 sc) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            typename RD::G g(threadIdx.x + q * TL::T);
            u32 v[1 << R];
            RD::load(lds, g, v);
            cm_dif_stages<R, RD::S0, RD::DLOG>(v, g.gi_lo, tw, tws, p);
#pragma unroll
            for (int m = 0; m < (1 << R); ++m)
                v[m] = cm_shoup(cm_mont(fr[q][m], v[m], p, pinv), sc.x, sc.y, p);
            cm_dit_stages<R, RD::S0, RD::DLOG>(v, g.gi_lo, twi, tws, p);
            RD::store(lds, g, v);
        }
        __syncthreads();
    }
};
