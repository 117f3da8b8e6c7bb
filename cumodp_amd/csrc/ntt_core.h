// ntt_core.h -- in-LDS number-theoretic transform building blocks for gfx950.
//
// A tile holds C independent length-L vectors (L = 2^LOGL) in LDS, either
// row-major (C rows of L contiguous words) or column-major (L rows of C
// words, i.e. C columns of a larger matrix).  Transforms run in "rounds":
// one round applies R consecutive radix-2 stages to groups of 2^R elements
// that each thread holds in registers, so a 4096-point transform is three
// register-resident radix-16 passes with two LDS exchanges in between.
// Rounds can run on NV tiles at once (the two operands of a product), so
// each twiddle is loaded once for both.
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
// blocks of 2d words, and maps (x, y) -> (x + y, (x - y) * w_{2d}^(q mod d)).
//
// Twiddles come from a level table: for every M = 2, 4, ..., 2^LT the M/2
// entries w_M^0 .. w_M^(M/2-1) (Montgomery forms w 2^32 mod p) are stored
// contiguously at offset M/2 - 1.  Stage s of an L-point transform reads
// level M = L >> s at index q mod d, so consecutive lanes read consecutive
// entries (one or a few cache lines per wave instruction).
#pragma once

#include "modarith.h"

#define CM_DEV __device__ __forceinline__

// One pad word after every 16 LDS words keeps the stride-16 access of the
// last round (and the column tiles) free of bank conflicts.
CM_DEV int cm_pad(int i) { return i + (i >> 4); }

CM_DEV u32 cm_brev(u32 x, int bits) { return __brev(x) >> (32 - bits); }

// Modulus and -p^-1 mod 2^32 for the lazy Montgomery products; every
// butterfly returns canonical residues ([0, p)).  Any odd p < 2^31.
struct Mod {
    u32 p, np;
    static constexpr bool LAZY = false;
};

// The same for p < 2^30 with values kept in [0, 2p) (Harvey's lazy
// butterflies, here on Montgomery products): a sum or difference of two such
// values is below 4p < 2^32 and is brought back below 2p with one
// conditional subtraction of 2p, and the Montgomery product of a value below
// 4p by a twiddle below p is below p 2^32, so it lands in [0, 2p) with no
// final subtraction.  Two instructions fewer per twiddled butterfly; results
// that leave a transform pass through cm_canon.
struct ModLazy {
    u32 p, np;
    static constexpr bool LAZY = true;
};

// The same modulus as a value of this basic block.  Code that follows a
// branch (the checked and unchecked first rounds) would otherwise take the
// 64-bit zero extension of p from the entry block, whose zero high word the
// instruction selector cannot see across blocks: it then expands each
// m * p + t of a Montgomery product into two v_mad_u64_u32 and a v_mov_b32
// instead of one v_mad_u64_u32.  readfirstlane of a uniform value costs no
// vector instruction.
template <class MD>
CM_DEV MD cm_rebind(MD md) {
    md.p = __builtin_amdgcn_readfirstlane(md.p);
    md.np = __builtin_amdgcn_readfirstlane(md.np);
    return md;
}

// the range bound c of the values between butterflies: p or 2p
template <class MD>
CM_DEV u32 cm_bound(MD md) { return MD::LAZY ? md.p << 1 : md.p; }

// a * w for a twiddle or constant w < p in Montgomery form and a < 2c:
// canonical for Mod, in [0, 2p) for ModLazy
template <class MD>
CM_DEV u32 cm_mulw(u32 a, u32 w, MD md) {
    if constexpr (MD::LAZY) return cm_mont_lazy(a, w, md.p, md.np);
    else return cm_mul_tw(a, w, md.p, md.np);
}

// canonical residue of a value in [0, c)
template <class MD>
CM_DEV u32 cm_canon(u32 a, MD md) {
    if constexpr (MD::LAZY) return cm_reduce2p(a, md.p);
    else return a;
}

// Twiddle w = tab[base + mm + lane] of a level table (base and mm uniform,
// lane per thread; lane == 0 when the index is uniform).  Plain loads: a raw
// buffer form (uniform part as the scalar offset) measured slower here,
// 80 against 75 us on the forward column pass of the 2^24 multiply.
template <bool UNIFORM>
CM_DEV u32 cm_twiddle(const u32 *tab, int base_mm, int lane) {
    if constexpr (UNIFORM) return tab[base_mm];
    else return tab[base_mm + lane];
}

// (x, y) -> (x + y, (x - y) w) on values in [0, c): the difference enters
// the product unreduced (x - y + c < 2c, and 2c w < p 2^32).  cm_reduce2p and
// cm_sub are exact for any bound c with 2c < 2^32.
//
// Canonical residues (Mod, any p < 2^31) take the difference signed instead:
// e = x - y lies in (-p, p) and t = e w in (-p^2, p^2).  With m = t (-p^-1)
// mod 2^32 read as a signed 32-bit value, t + m p is divisible by 2^32, and
// r = (t + m p) / 2^32 = e w / 2^32 mod p has |r| < (p^2 + 2^31 p) / 2^32 < p.
// One conditional add of p (the smaller of r and r + p as unsigned words)
// makes it canonical.  That is a subtraction, two v_mad_i64_i32, a
// v_mul_lo_u32, an add and a min: one instruction fewer than adding p before
// an unsigned Montgomery product and reducing after.
CM_DEV u32 cm_mul_signed(u32 e, u32 w, u32 p, u32 np) {
    const int64_t t = (int64_t)(int32_t)e * (int32_t)w;
    const int32_t m = (int32_t)((u32)t * np);
    const u32 r = (u32)(((int64_t)m * (int32_t)p + t) >> 32);
    const u32 rp = r + p;
    return rp < r ? rp : r;
}

template <class MD>
CM_DEV void cm_dif_bfly(u32 &x, u32 &y, u32 w, MD md) {
    const u32 c = cm_bound(md);
    u32 s = cm_reduce2p(x + y, c);
    if constexpr (MD::LAZY) {
        u32 d = x - y + c;             // (0, 2c)
        y = cm_mulw(d, w, md);
    } else {
        y = cm_mul_signed(x - y, w, md.p, md.np);
    }
    x = s;
}

template <class MD>
CM_DEV void cm_dif_bfly1(u32 &x, u32 &y, MD md) {
    const u32 c = cm_bound(md);
    u32 s = cm_reduce2p(x + y, c);
    y = cm_sub(x, y, c);
    x = s;
}

// (a, b) -> (a + b w, a - b w)
template <class MD>
CM_DEV void cm_dit_bfly(u32 &a, u32 &b, u32 w, MD md) {
    const u32 c = cm_bound(md);
    u32 t = cm_mulw(b, w, md);
    b = cm_sub(a, t, c);
    a = cm_reduce2p(a + t, c);
}

template <class MD>
CM_DEV void cm_dit_bfly1(u32 &a, u32 &b, MD md) {
    const u32 c = cm_bound(md);
    u32 t = b;
    b = cm_sub(a, t, c);
    a = cm_reduce2p(a + t, c);
}

template <int LOGL_, int C_, bool COLS_, int T_ = (C_ << LOGL_) / 16>
struct Tile {
    static constexpr int LOGL = LOGL_;
    static constexpr int L = 1 << LOGL_;
    static constexpr int C = C_;
    static constexpr bool COLS = COLS_;
    static constexpr int WORDS = C_ * L;
    // LDS placement of tile-linear index x.  Rows of 2^13 words and more XOR
    // the three bits above each 64-word block into the low six
    // (x ^ 9 ((x >> 6) & 7)): every round's wave access and the linear load
    // and store phases then hit 64 distinct banks (at worst two per bank on a
    // 32-bank LDS), with no padding; the row pass's bank-conflict cycles drop
    // eightfold.
    // Column tiles of 4 or more columns pad every 64-word block by K = 8 words
    // (4 for four columns): x + K (x >> 6).  A wave's 64 lanes then hit 64
    // distinct banks in every round (the last round's lanes stride 64 words,
    // one block apart, and land K banks apart), a four-word group stays
    // contiguous and 16-byte aligned for the 128-bit load and store phases,
    // and each thread's elements of a round sit at constant offsets from its
    // first (the XOR swizzle removes the conflicts too, but its address
    // arithmetic does not fold into LDS instruction offsets and pushes the
    // column kernels past 64 VGPRs).  Other tiles keep one pad word per 16.
    static constexpr bool SWZ = !COLS_ && LOGL_ >= 13;
    static constexpr int BLKPAD = (COLS_ && C_ >= 4) ? (C_ >= 8 ? 8 : 4) : 0;
    static constexpr int LDS_WORDS =
        SWZ ? WORDS : (BLKPAD ? WORDS + BLKPAD * (WORDS >> 6) : WORDS + (WORDS >> 4));
    static CM_DEV int phys(int x) {
        if constexpr (SWZ) return x ^ (((x >> 6) & 7) * 9);
        else if constexpr (BLKPAD != 0) return x + BLKPAD * (x >> 6);
        else return cm_pad(x);
    }
    // four consecutive words x .. x+3 (x a multiple of 4) are contiguous and
    // 16-byte aligned in LDS
    static constexpr bool QUAD = BLKPAD != 0;
    static constexpr int T = T_;                   // threads per workgroup
    static constexpr int NR = (LOGL_ + 3) / 4;     // rounds
    static constexpr int rsz(int j) { return LOGL_ / NR + (j < LOGL_ % NR ? 1 : 0); }
    static constexpr int s0(int j) {
        int s = 0;
        for (int i = 0; i < j; ++i) s += rsz(i);
        return s;
    }
    // LDS address of element i of vector c
    static CM_DEV int addr(int c, int i) { return phys(COLS_ ? i * C_ + c : c * L + i); }
    // group g of a round with 2^(LOGL-R) groups per vector -> (vector, index)
    template <int R>
    static CM_DEV void split(int g, int &c, int &gi) {
        if (COLS_) { c = g % C_; gi = g / C_; }
        else { c = g >> (LOGL_ - R); gi = g & ((1 << (LOGL_ - R)) - 1); }
    }
};

// Words idx .. idx+3 of a tile (idx a multiple of 4): one 128-bit LDS access
// when the layout keeps them together (the LDS array is 16-byte aligned).
template <class TL>
CM_DEV void cm_lds_put4(u32 *lds, int idx, uint4 v) {
    if constexpr (TL::QUAD) {
        *(uint4 *)(lds + TL::phys(idx)) = v;
    } else {
        lds[TL::phys(idx)] = v.x;
        lds[TL::phys(idx + 1)] = v.y;
        lds[TL::phys(idx + 2)] = v.z;
        lds[TL::phys(idx + 3)] = v.w;
    }
}

template <class TL>
CM_DEV uint4 cm_lds_get4(const u32 *lds, int idx) {
    if constexpr (TL::QUAD) {
        return *(const uint4 *)(lds + TL::phys(idx));
    } else {
        uint4 v;
        v.x = lds[TL::phys(idx)];
        v.y = lds[TL::phys(idx + 1)];
        v.z = lds[TL::phys(idx + 2)];
        v.w = lds[TL::phys(idx + 3)];
        return v;
    }
}

// R DIF stages (S0 .. S0+R-1 of an L = 2^LOGL transform) on NV register
// groups that share their positions (and so their twiddles).  TOPZERO: the
// upper half of the group (m >= 2^(R-1)) holds zeros, so the first stage is
// (x, 0) -> (x, x w) and those inputs are never read.
template <int LOGL, int R, int S0, int DLOG, int NV, bool TOPZERO = false, class MD>
CM_DEV void cm_dif_stages(u32 (&v)[NV][1 << R], int gi_lo, const u32 *__restrict__ lv, MD md) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int h = 1 << (R - 1 - t);
        const int base = (1 << (LOGL - S0 - t - 1)) - 1;     // level L >> (S0+t)
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) {
            if (m & h) continue;
            const int mm = (m & (h - 1)) << DLOG;
            if (TOPZERO && t == 0) {
                if (DLOG == 0 && mm == 0) {
#pragma unroll
                    for (int a = 0; a < NV; ++a) v[a][m + h] = v[a][m];
                } else {
                    const u32 w = cm_twiddle<DLOG == 0>(lv, base + mm, gi_lo);
#pragma unroll
                    for (int a = 0; a < NV; ++a) v[a][m + h] = cm_mulw(v[a][m], w, md);
                }
                continue;
            }
            if (DLOG == 0 && mm == 0) {
#pragma unroll
                for (int a = 0; a < NV; ++a) cm_dif_bfly1(v[a][m], v[a][m + h], md);
            } else {
                const u32 w = cm_twiddle<DLOG == 0>(lv, base + mm, gi_lo);
#pragma unroll
                for (int a = 0; a < NV; ++a) cm_dif_bfly(v[a][m], v[a][m + h], w, md);
            }
        }
    }
}

// Inverse of cm_dif_stages with the inverse level table.
template <int LOGL, int R, int S0, int DLOG, int NV, class MD>
CM_DEV void cm_dit_stages(u32 (&v)[NV][1 << R], int gi_lo, const u32 *__restrict__ lvi, MD md) {
#pragma unroll
    for (int t = R - 1; t >= 0; --t) {
        const int h = 1 << (R - 1 - t);
        const int base = (1 << (LOGL - S0 - t - 1)) - 1;
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) {
            if (m & h) continue;
            const int mm = (m & (h - 1)) << DLOG;
            if (DLOG == 0 && mm == 0) {
#pragma unroll
                for (int a = 0; a < NV; ++a) cm_dit_bfly1(v[a][m], v[a][m + h], md);
            } else {
                const u32 w = cm_twiddle<DLOG == 0>(lvi, base + mm, gi_lo);
#pragma unroll
                for (int a = 0; a < NV; ++a) cm_dit_bfly(v[a][m], v[a][m + h], w, md);
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

    template <int NV>
    static CM_DEV void load(u32 *const (&lds)[NV], const G &g, u32 (&v)[NV][1 << R]) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int m = 0; m < (1 << R); ++m) v[a][m] = lds[a][g.a(m)];
    }
    template <int NV>
    static CM_DEV void store(u32 *const (&lds)[NV], const G &g, const u32 (&v)[NV][1 << R]) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int m = 0; m < (1 << R); ++m) lds[a][g.a(m)] = v[a][m];
    }
    // full forward round on NV tiles, LDS -> LDS; TOPZERO (first round
    // only): the upper half of every vector is zero and is not read
    template <int NV, bool TOPZERO = false, class MD>
    static CM_DEV void dif(u32 *const (&lds)[NV], const u32 *lv, MD md) {
        static_assert(!TOPZERO || S0 == 0, "zero upper half only in the first round");
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            G g(threadIdx.x + q * TL::T);
            u32 v[NV][1 << R];
            if (TOPZERO) {
#pragma unroll
                for (int a = 0; a < NV; ++a)
#pragma unroll
                    for (int m = 0; m < (1 << (R - 1)); ++m) v[a][m] = lds[a][g.a(m)];
            } else {
                load<NV>(lds, g, v);
            }
            cm_dif_stages<TL::LOGL, R, S0, DLOG, NV, TOPZERO>(v, g.gi_lo, lv, md);
            store<NV>(lds, g, v);
        }
    }
    // full inverse round on NV tiles, LDS -> LDS
    template <int NV, class MD>
    static CM_DEV void dit(u32 *const (&lds)[NV], const u32 *lvi, MD md) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            G g(threadIdx.x + q * TL::T);
            u32 v[NV][1 << R];
            load<NV>(lds, g, v);
            cm_dit_stages<TL::LOGL, R, S0, DLOG, NV>(v, g.gi_lo, lvi, md);
            store<NV>(lds, g, v);
        }
    }
};

// Forward rounds [J, J1) on NV tiles with a barrier after each.
template <class TL, int J, int J1, int NV, class MD>
CM_DEV void cm_dif_rounds(u32 *const (&lds)[NV], const u32 *lv, MD md) {
    if constexpr (J < J1) {
        Round<TL, J>::template dif<NV>(lds, lv, md);
        __syncthreads();
        cm_dif_rounds<TL, J + 1, J1, NV>(lds, lv, md);
    }
}

// Inverse rounds J, J-1, ..., J0 (J >= J0) with a barrier after each.
template <class TL, int J, int J0, int NV, class MD>
CM_DEV void cm_dit_rounds(u32 *const (&lds)[NV], const u32 *lvi, MD md) {
    if constexpr (J >= J0) {
        Round<TL, J>::template dit<NV>(lds, lvi, md);
        __syncthreads();
        cm_dit_rounds<TL, J - 1, J0, NV>(lds, lvi, md);
    }
}

// Full forward transform of every vector of one tile whose upper halves are
// zero (and were not stored to LDS).
template <class TL, class MD>
CM_DEV void cm_tile_dif_topzero(u32 *lds, const u32 *lv, MD md) {
    u32 *const t[1] = {lds};
    Round<TL, 0>::template dif<1, true>(t, lv, md);
    __syncthreads();
    cm_dif_rounds<TL, 1, TL::NR, 1>(t, lv, md);
}

// Full forward transform of every vector of one tile (bit-reversed result).
template <class TL, class MD>
CM_DEV void cm_tile_dif(u32 *lds, const u32 *lv, MD md) {
    u32 *const t[1] = {lds};
    cm_dif_rounds<TL, 0, TL::NR, 1>(t, lv, md);
}

template <class TL, class MD>
CM_DEV void cm_tile_dit(u32 *lds, const u32 *lvi, MD md) {
    u32 *const t[1] = {lds};
    cm_dit_rounds<TL, TL::NR - 1, 0, 1>(t, lvi, md);
}

// Column tiles (COLS) whose first and last rounds talk to global memory
// directly: element (row, c) of the tile is X[row n2 + c], X at the tile's
// first column.  Each thread's round-group is one column, so a wave reads or
// writes C-column row segments; the row-dependent part of an address is
// uniform (m 2^DLOG n2), which keeps the per-lane offset in one 32-bit
// register.  No LDS staging pass and no barrier precede the first round or
// follow the last.

// Word `off` of a row whose base is uniform, addressed as base + 4 off with a
// 32-bit byte offset.  off < 2^30 (tiles of transforms up to 2^26 words).
CM_DEV u32 cm_ld_row(const u32 *base, u32 off) {
    return *(const u32 *)((const char *)base + (off << 2));
}
CM_DEV void cm_st_row(u32 *base, u32 off, u32 v) { *(u32 *)((char *)base + (off << 2)) = v; }

// Raw buffer view of a tile: the tile base in a scalar resource, each access
// = per-lane byte offset (vector register) + uniform row byte offset (scalar
// register), so no vector instruction computes an address.  No range limit
// (every access stays inside the arrays the caller passed); word 3 is the
// raw-buffer format of the gfx9 family.
CM_DEV __amdgpu_buffer_rsrc_t cm_tile_rsrc(const u32 *base) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, -1, 0x00020000);
}
CM_DEV u32 cm_buf_ld(__amdgpu_buffer_rsrc_t r, u32 off_bytes, u32 row_bytes) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off_bytes, (int)row_bytes, 0);
}
CM_DEV void cm_buf_st(__amdgpu_buffer_rsrc_t r, u32 off_bytes, u32 row_bytes, u32 v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)off_bytes, (int)row_bytes, 0);
}

// Forward round 0 from global memory into LDS.  CHECK: elements at or beyond
// lim (relative to X) read as zero.  TOPZERO: rows >= L/2 are zero and not
// read.
template <class TL, bool TOPZERO, bool CHECK, class MD>
CM_DEV void cm_col_dif_first(u32 *lds, const u32 *X, int64_t n2, int64_t lim, const u32 *lv,
                             MD md) {
    using RD = Round<TL, 0>;
    constexpr int R = RD::R, M = TOPZERO ? (1 << (R - 1)) : (1 << R);
#pragma unroll
    for (int q = 0; q < RD::GPT; ++q) {
        typename RD::G g(threadIdx.x + q * TL::T);
        const u32 off = (u32)g.i0 * (u32)n2 + (u32)g.c;
        u32 v[1][1 << R];
        if constexpr (CHECK) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int64_t rowoff = (int64_t)(m << RD::DLOG) * n2;
                v[0][m] = rowoff + off < lim ? cm_ld_row(X + rowoff, off) : 0u;
            }
        } else {
            const auto rx = cm_tile_rsrc(X);
#pragma unroll
            for (int m = 0; m < M; ++m)
                v[0][m] = cm_buf_ld(rx, off << 2, (u32)((m << RD::DLOG) * n2) << 2);
        }
        cm_dif_stages<TL::LOGL, R, RD::S0, RD::DLOG, 1, TOPZERO>(v, g.gi_lo, lv, md);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) lds[g.a(m)] = v[0][m];
    }
}

// Last forward round from LDS straight to global memory (all rows).
template <class TL, class MD>
CM_DEV void cm_col_dif_last(const u32 *lds, u32 *Y, int64_t n2, const u32 *lv, MD md) {
    using RD = Round<TL, TL::NR - 1>;
    constexpr int R = RD::R;
    const auto ry = cm_tile_rsrc(Y);
#pragma unroll
    for (int q = 0; q < RD::GPT; ++q) {
        typename RD::G g(threadIdx.x + q * TL::T);
        const u32 off = (u32)g.i0 * (u32)n2 + (u32)g.c;
        u32 v[1][1 << R];
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) v[0][m] = lds[g.a(m)];
        cm_dif_stages<TL::LOGL, R, RD::S0, RD::DLOG, 1>(v, g.gi_lo, lv, md);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m)
            cm_buf_st(ry, off << 2, (u32)((m << RD::DLOG) * n2) << 2, v[0][m]);
    }
}

// First inverse round (round NR-1) from global memory into LDS (all rows).
template <class TL, class MD>
CM_DEV void cm_col_dit_first(u32 *lds, const u32 *X, int64_t n2, const u32 *lvi, MD md) {
    using RD = Round<TL, TL::NR - 1>;
    constexpr int R = RD::R;
    const auto rx = cm_tile_rsrc(X);
#pragma unroll
    for (int q = 0; q < RD::GPT; ++q) {
        typename RD::G g(threadIdx.x + q * TL::T);
        const u32 off = (u32)g.i0 * (u32)n2 + (u32)g.c;
        u32 v[1][1 << R];
#pragma unroll
        for (int m = 0; m < (1 << R); ++m)
            v[0][m] = cm_buf_ld(rx, off << 2, (u32)((m << RD::DLOG) * n2) << 2);
        cm_dit_stages<TL::LOGL, R, RD::S0, RD::DLOG, 1>(v, g.gi_lo, lvi, md);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) lds[g.a(m)] = v[0][m];
    }
}

// Last inverse round (round 0) from LDS straight to global memory, canonical
// residues out.  CHECK: elements at or beyond lim (relative to Y) are not
// stored.
template <class TL, bool CHECK, class MD>
CM_DEV void cm_col_dit_last(const u32 *lds, u32 *Y, int64_t n2, int64_t lim, const u32 *lvi,
                            MD md) {
    using RD = Round<TL, 0>;
    constexpr int R = RD::R;
#pragma unroll
    for (int q = 0; q < RD::GPT; ++q) {
        typename RD::G g(threadIdx.x + q * TL::T);
        const u32 off = (u32)g.i0 * (u32)n2 + (u32)g.c;
        u32 v[1][1 << R];
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) v[0][m] = lds[g.a(m)];
        cm_dit_stages<TL::LOGL, R, RD::S0, RD::DLOG, 1>(v, g.gi_lo, lvi, md);
        if constexpr (CHECK) {
#pragma unroll
            for (int m = 0; m < (1 << R); ++m) {
                const int64_t rowoff = (int64_t)(m << RD::DLOG) * n2;
                if (rowoff + off < lim) cm_st_row(Y + rowoff, off, cm_canon(v[0][m], md));
            }
        } else {
            const auto ry = cm_tile_rsrc(Y);
#pragma unroll
            for (int m = 0; m < (1 << R); ++m)
                cm_buf_st(ry, off << 2, (u32)((m << RD::DLOG) * n2) << 2, cm_canon(v[0][m], md));
        }
    }
}

// Multiply of two tiles A (first operand, becomes the product) and B:
// forward transforms of both (shared twiddles), pointwise product scaled by
// sc (Montgomery form of 2^32 / N mod p: the two lazy Montgomery products
// divide by 2^64, sc restores 2^32 and supplies the inverse transform's 1/N)
// fused with the last forward and first inverse round, then the remaining
// inverse rounds on A.  With sc == 0 the product is left as F G / 2^32 (one
// Montgomery product) for a caller that applies 2^32 / N later.  Barriers
// included; the product is in A afterwards.
// The middle of a tile multiply: the last forward round of A and B, the
// pointwise product (scaled by sc as in cm_tile_multiply) and the first
// inverse round, LDS to LDS into A, then a barrier.
template <class TL, class MD>
CM_DEV void cm_tile_pointwise_round(u32 *A, u32 *B, const u32 *lv, const u32 *lvi, MD md,
                                    u32 sc) {
    u32 *const ab[2] = {A, B};
    using RD = Round<TL, TL::NR - 1>;
    constexpr int R = RD::R;
#pragma unroll
    for (int q = 0; q < RD::GPT; ++q) {
        typename RD::G g(threadIdx.x + q * TL::T);
        u32 v[2][1 << R];
        RD::template load<2>(ab, g, v);
        cm_dif_stages<TL::LOGL, R, RD::S0, RD::DLOG, 2>(v, g.gi_lo, lv, md);
        u32 w[1][1 << R];
#pragma unroll
        for (int m = 0; m < (1 << R); ++m)
            w[0][m] = sc ? cm_mulw(cm_mont_lazy(v[0][m], v[1][m], md.p, md.np), sc, md)
                         : cm_mulw(v[0][m], v[1][m], md);
        cm_dit_stages<TL::LOGL, R, RD::S0, RD::DLOG, 1>(w, g.gi_lo, lvi, md);
#pragma unroll
        for (int m = 0; m < (1 << R); ++m) A[g.a(m)] = w[0][m];
    }
    __syncthreads();
}

template <class TL, class MD>
CM_DEV void cm_tile_multiply(u32 *A, u32 *B, const u32 *lv, const u32 *lvi, MD md, u32 sc) {
    u32 *const ab[2] = {A, B};
    cm_dif_rounds<TL, 0, TL::NR - 1, 2>(ab, lv, md);
    cm_tile_pointwise_round<TL>(A, B, lv, lvi, md, sc);
    u32 *const a1[1] = {A};
    cm_dit_rounds<TL, TL::NR - 2, 0, 1>(a1, lvi, md);
}

// The same multiply for a tile of row vectors (16 elements per thread, first
// round radix-16) whose round-0 groups are in registers on entry (fa for A,
// ga for B) and whose product's round-0 groups are left in fa: the first
// forward and the last inverse round skip the LDS staging pass and its
// barrier.  Thread t's group is vector c = t >> (LOGL-4), elements
// gi + m 2^(LOGL-4), gi = t mod 2^(LOGL-4) (elements t + T m when C == 1).
// sc as in cm_tile_multiply.
template <class TL, class MD>
CM_DEV void cm_row_multiply(u32 *A, u32 *B, u32 (&fa)[16], const u32 (&ga)[16], const u32 *lv,
                            const u32 *lvi, MD md, u32 sc) {
    using R0 = Round<TL, 0>;
    static_assert(!TL::COLS && R0::R == 4 && R0::GPT == 1 && TL::NR >= 2,
                  "row tiles of 16 elements per thread, radix-16 first round");
    const typename R0::G g(threadIdx.x);
    u32 v[2][16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        v[0][m] = fa[m];
        v[1][m] = ga[m];
    }
    cm_dif_stages<TL::LOGL, 4, 0, R0::DLOG, 2>(v, g.gi_lo, lv, md);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        A[g.a(m)] = v[0][m];
        B[g.a(m)] = v[1][m];
    }
    __syncthreads();
    u32 *const ab[2] = {A, B};
    cm_dif_rounds<TL, 1, TL::NR - 1, 2>(ab, lv, md);
    cm_tile_pointwise_round<TL>(A, B, lv, lvi, md, sc);
    u32 *const a1[1] = {A};
    cm_dit_rounds<TL, TL::NR - 2, 1, 1>(a1, lvi, md);
    u32 w[1][16];
#pragma unroll
    for (int m = 0; m < 16; ++m) w[0][m] = A[g.a(m)];
    cm_dit_stages<TL::LOGL, 4, 0, R0::DLOG, 1>(w, g.gi_lo, lvi, md);
#pragma unroll
    for (int m = 0; m < 16; ++m) fa[m] = w[0][m];
}
