// batch.hip -- one call multiplying many pairs of uneven degree mod p.
//
// The reference has no list multiply for uneven degrees: list_stockham_dev
// (src/list_stockham.cu:595-630) transforms m rows of one length, and the
// unbuilt draft src/.list_fft_poly_mul.cu:15-53 multiplies equal-length
// pairs.  Here products are binned by transform size 2^e:
//   * 2^e <= W (W = 2048, 4096 or 8192 words, a tuned choice): one fused
//     launch for every such product.  Each workgroup takes a W-word tile
//     holding W / 2^e products of one size and runs the whole multiply in
//     LDS; the tile's size selects the code path at run time, so one launch
//     covers all sizes;
//   * 2^e <= 2^13 above W: one launch of the 8192-point workgroup multiply;
//   * e >= 14: per size, the inputs are gathered into zero-padded slabs,
//     multiplied by the batched four-step kernels and scattered back.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "accum.h"
#include "internal.h"
#include "ntt_core.h"

namespace cm {

struct TileDesc {
    int32_t first;   // first descriptor of the tile
    int32_t count;   // products in the tile (<= 4096 >> logn)
    int32_t logn;
    int32_t pad;
};

// Mixed-size tile variants (words per tile, 16 per thread); the
// tuner picks one per process (cumodp_tune_set_list_tile).  Products of
// 2^logn <= words points go to the mixed kernel, words / 2^logn per tile.
constexpr int kMixVariants = 3;
constexpr int kMixWordsOf[kMixVariants] = {2048, 4096, 8192};
int g_mix_variant = 2;   // 8192 words, 512 threads: measured default (4096 uneven pairs)

struct Scales {
    u32 v[16];
};

template <int LOGL, int W, int T>
using MixTile = Tile<LOGL, (W >> LOGL), false, T>;

template <int LOGL, int W, int T, class MD>
CM_DEV void mul_tile(u32 *A, u32 *B, const u32 *F, const u32 *G, u32 *H, const MulDesc *d,
                     int cnt, const u32 *lv, const u32 *lvi, MD md, u32 sc) {
    using TL = MixTile<LOGL, W, T>;
    constexpr int L = TL::L;
    constexpr int kMixT = T;
    // element idx of the tile: vector c = idx >> LOGL, position i; all loads
    // of a thread are issued before the first LDS store
    constexpr int PER = TL::WORDS / kMixT;
    u32 fa[PER], ga[PER];
    if constexpr (Round<TL, 0>::R == 4 && Round<TL, 0>::GPT == 1 && TL::NR >= 2) {
        // thread t loads its group of the first forward and the last inverse
        // round (product c, elements gi + r 2^(LOGL-4)), which run on the
        // registers
        constexpr int S = LOGL - 4;
        const int c = threadIdx.x >> S, gi = threadIdx.x & ((1 << S) - 1);
        const bool live = c < cnt;
        const MulDesc dc = live ? d[c] : MulDesc{};
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = gi + (r << S);
            fa[r] = live && i < dc.lf ? F[dc.f + i] : 0u;
            ga[r] = live && i < dc.lg ? G[dc.g + i] : 0u;
        }
        cm_row_multiply<TL>(A, B, fa, ga, lv, lvi, md, sc);
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = gi + (r << S);
            if (live && i < dc.lh) H[dc.h + i] = cm_canon(fa[r], md);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * kMixT, c = idx >> LOGL, i = idx & (L - 1);
        u32 f = 0, g = 0;
        if (c < cnt) {
            if (i < d[c].lf) f = F[d[c].f + i];
            if (i < d[c].lg) g = G[d[c].g + i];
        }
        fa[r] = f;
        ga[r] = g;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + r * kMixT, c = idx >> LOGL, i = idx & (L - 1);
        A[TL::addr(c, i)] = fa[r];
        B[TL::addr(c, i)] = ga[r];
    }
    __syncthreads();
    cm_tile_multiply<TL>(A, B, lv, lvi, md, sc);
    for (int idx = threadIdx.x; idx < TL::WORDS; idx += kMixT) {
        const int c = idx >> LOGL, i = idx & (L - 1);
        if (c < cnt && i < d[c].lh) H[d[c].h + i] = cm_canon(A[TL::addr(c, i)], md);
    }
}

template <int W, int T, class MD>
__global__ __launch_bounds__(T) void k_mul_mixed(
    const u32 *__restrict__ F, const u32 *__restrict__ G, u32 *__restrict__ H,
    const MulDesc *__restrict__ descs, const TileDesc *__restrict__ tiles,
    const u32 *__restrict__ lv, const u32 *__restrict__ lvi, MD md, Scales scales) {
    constexpr int LW = W + (W >> 4);
    __shared__ u32 lds[2 * LW];
    u32 *A = lds, *B = lds + LW;
    const TileDesc t = tiles[blockIdx.x];
    const MulDesc *d = descs + t.first;
    const u32 sc = scales.v[t.logn];
#define CM_MIX(K)                                                                        \
    case K:                                                                              \
        if constexpr ((1 << K) <= W && (W >> K) * (1 << K) / T >= 16)                    \
            mul_tile<K, W, T>(A, B, F, G, H, d, t.count, lv, lvi, md, sc);               \
        break;
    switch (t.logn) {
        CM_MIX(4) CM_MIX(5) CM_MIX(6) CM_MIX(7) CM_MIX(8) CM_MIX(9) CM_MIX(10) CM_MIX(11)
        CM_MIX(12) CM_MIX(13)
    default: break;
    }
#undef CM_MIX
}

// zero-padded copy of operand `which` (0: F, 1: G) of descriptor row y into
// row y of a slab of n-word rows
// Row y of the slab (stride n words) gets the first m words of operand y,
// zero-padded: m is the longest operand of the batch, and the column pass
// reads no further.
__global__ void k_gather(u32 *__restrict__ slab, int64_t n, int64_t m,
                         const u32 *__restrict__ src, const MulDesc *__restrict__ d, int which) {
    const MulDesc dd = d[blockIdx.y];
    const int64_t off = which ? dd.g : dd.f;
    const int len = which ? dd.lg : dd.lf;
    u32 *row = slab + (int64_t)blockIdx.y * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        row[i] = i < len ? src[off + i] : 0u;
}

__global__ void k_scatter(u32 *__restrict__ H, const u32 *__restrict__ slab, int64_t n,
                          const MulDesc *__restrict__ d) {
    const MulDesc dd = d[blockIdx.y];
    const u32 *row = slab + (int64_t)blockIdx.y * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dd.lh;
         i += (int64_t)gridDim.x * blockDim.x)
        H[dd.h + i] = row[i];
}

// ---- one launch for the whole list ------------------------------------------
// One launch for every product of a list call up to 2^15 points, schoolbook
// and NTT alike: one workgroup per task (1024 threads, about 136 KiB of LDS,
// so one per CU at a time).  Each workgroup takes its task from a list in
// descending order of estimated cost through an atomic counter, so the tasks
// start largest first whatever order the hardware dispatches workgroups in:
//   * tile: 16384 words of products of one size 2^logn <= 2^14 (16384 >>
//     logn products; one 2^14 product), multiplied in LDS as by k_mul_mixed;
//   * big half: one of the two 2^14-point halves (cyclic, negacyclic) of a
//     2^15-point product, recombined by a short kernel after the launch;
//   * plain: 4096 consecutive outputs of one schoolbook product (16 waves of
//     256 outputs, each wave on its own, the arithmetic of k_plain_mul).
// Largest tasks first keeps the tail short; the counter starts at zero in
// every call's uploaded descriptor blob.  (A loop over tasks in fewer
// workgroups was measured slower: inlined into the loop, the task bodies'
// thread-index arithmetic is hoisted out of it and spills, and called out of
// line they take their uniform pointers in vector registers.)
constexpr int kPT = 1024;                    // threads of the persistent workgroup
constexpr int kPW = 16384;                   // words per product tile
constexpr int kPLW = kPW + (kPW >> 4);       // LDS words of one operand tile (padded layouts)
constexpr int kPlainTaskOut = 4096;          // outputs of a plain task (16 waves x 256)
constexpr int kPlainJT = 512;                // B coefficients per pass of a plain task

enum : int32_t { kTaskTile = 0, kTaskBigCyc = 1, kTaskBigNeg = 2, kTaskPlain = 3 };

struct ListTask {
    int32_t kind;
    int32_t first;   // tile: first descriptor; big half, plain: the descriptor
    int32_t count;   // tile: products; big half: the product's scratch slot;
                     // plain: first output of the task
    int32_t logn;    // tile: transform size
};

// One 2^15-point product as two 2^14-point products in LDS (n = 2^14, the
// F and G tiles of a 2^15-point product would not fit together):
//   A = H mod x^n - 1, the cyclic product of F_lo + F_hi and G_lo + G_hi;
//   B = H mod x^n + 1, the negacyclic product of F_lo - F_hi and G_lo - G_hi,
//       taken as the cyclic product of the operands twisted by psi^i (psi a
//       primitive 2^15-th root, psi^n = -1) and untwisted by psi^-i.
// With the pointwise constant of a 2^15-point product both come out halved,
// so H_i = A_i/2 + psi^-i B'_i/2 and H_(n+i) = A_i/2 - psi^-i B'_i/2.  psi^i
// and psi^-i are entries of the 2^15 level tables.
// The halves are separate tasks: the cyclic one leaves A/2 in H_0 ..
// H_(n-1) (inside the output: the product has more than n coefficients), the
// negacyclic one psi^-i B'/2 in the product's slot of the scratch S, and
// k_big_combine adds and subtracts them after the launch.  (Both halves in
// one workgroup spill: the operand loads of the second are scheduled beside
// the registers of the first.)  Thread t holds elements i = t + 1024 r.
template <bool NEGA, class MD>
CM_DEV void task_big_half(u32 *lds, const u32 *F, const u32 *G, u32 *H, const MulDesc *dp,
                          u32 *S, int slot, const u32 *lv, const u32 *lvi, MD md, u32 sc15) {
    using TL = Tile<14, 1, false, kPT>;
    constexpr int n = 1 << 14;
    const MulDesc d = *dp;
    const u32 p = md.p;
    const int t = threadIdx.x;
    u32 fa[16], ga[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = t + (r << 10);
        const u32 f0 = i < d.lf ? F[d.f + i] : 0u, f1 = n + i < d.lf ? F[d.f + n + i] : 0u;
        const u32 g0 = i < d.lg ? G[d.g + i] : 0u, g1 = n + i < d.lg ? G[d.g + n + i] : 0u;
        if constexpr (NEGA) {
            const u32 psi = lv[n - 1 + i];
            fa[r] = cm_mulw(cm_sub(f0, f1, p), psi, md);
            ga[r] = cm_mulw(cm_sub(g0, g1, p), psi, md);
        } else {
            fa[r] = cm_add(f0, f1, p);
            ga[r] = cm_add(g0, g1, p);
        }
    }
    cm_row_multiply<TL>(lds, lds + n, fa, ga, lv, lvi, md, sc15);
    u32 *Sp = S + (int64_t)slot * n;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = t + (r << 10);
        if constexpr (NEGA) Sp[i] = cm_mulw(fa[r], lvi[n - 1 + i], md);
        else H[d.h + i] = fa[r];
    }
}

// H_i = A_i/2 + S_i, H_(n+i) = A_i/2 - S_i for the 2^15-point products
// (descriptors descs[0 .. count), slot j at S + j n), canonical out.  A/2
// and S are in [0, c), c = p or 2p.
template <class MD>
__global__ void k_big_combine(u32 *__restrict__ H, const MulDesc *__restrict__ descs,
                              const u32 *__restrict__ S, MD md) {
    constexpr int n = 1 << 14;
    const MulDesc d = descs[blockIdx.y];
    const u32 c = cm_bound(md);
    const u32 *Sp = S + (int64_t)blockIdx.y * n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u32 a = H[d.h + i], u = Sp[i];
        H[d.h + i] = cm_canon(cm_reduce2p(a + u, c), md);
        if (n + i < d.lh) H[d.h + n + i] = cm_canon(cm_sub(a, u, c), md);
    }
}

// Outputs c0 .. c0 + 4095 of the schoolbook product d: wave w computes the
// 256 outputs from c0 + 256 w on its own (its own LDS region, no workgroup
// barrier; a wave past the end of the product leaves at once), lane t four of
// them, with the arithmetic of k_plain_mul.  LDS operations of one wavefront
// complete in order, so a wave reads back what its lanes just stored.
constexpr int kPlainWaveWords = 256 + kPlainJT + 8 + kPlainJT;   // A window + B slice

CM_DEV void task_plain(u32 *lds, const u32 *F, const u32 *G, u32 *H, const MulDesc *dp, int c0,
                       PlainMod m) {
    const MulDesc d = *dp;
    const bool sw = d.logn != 0;
    const u32 *A = sw ? G + d.g : F + d.f;
    const u32 *B = sw ? F + d.f : G + d.g;
    const int la = sw ? d.lg : d.lf, lb = sw ? d.lf : d.lg;
    const int t = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cw = c0 + 256 * wv;
    if (cw >= d.lh) return;
    u32 *sa = lds + wv * kPlainWaveWords, *sb = sa + 256 + kPlainJT + 8;
    // terms j of output c need 0 <= c - j < la and 0 <= j < lb
    int j_lo = cw - (la - 1);
    j_lo = j_lo < 0 ? 0 : j_lo & ~3;
    int j_hi = cw + 255;
    j_hi = j_hi < lb - 1 ? j_hi : lb - 1;
    u64 acc[4] = {0, 0, 0, 0};
    for (int js = j_lo; js <= j_hi; js += kPlainJT) {
        const int nj = (j_hi - js + 1 + 3) & ~3;
        const int jn = nj < kPlainJT ? nj : kPlainJT;
        const int je = js + jn;
        // sa[x] = A[cw - je + x] (zero outside [0, la)), x < 256 + jn;
        // sb[k] = B[js + k] (zero beyond lb)
        const int a_lo = cw - je;
        __builtin_amdgcn_wave_barrier();
        for (int x = t; x < 256 + jn; x += 64) {
            const int i = a_lo + x;
            sa[x] = (i >= 0 && i < la) ? A[i] : 0u;
        }
        for (int k = t; k < jn; k += 64) {
            const int j = js + k;
            sb[k] = j < lb ? B[j] : 0u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const u32 *sw4 = sa + 4 * t;
        uint4 w = *(const uint4 *)(sw4 + jn);
        for (int k = 0; k < jn; k += 4) {
            const uint4 nw = *(const uint4 *)(sw4 + jn - k - 4);
            const uint4 bv = *(const uint4 *)(sb + k);
            const u32 x[8] = {nw.x, nw.y, nw.z, nw.w, w.x, w.y, w.z, w.w};
            const u32 b[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] += (u64)x[4 + r - s] * b[s];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = plain_fold(acc[r], m.c32);
            w = nw;
        }
    }
    u32 *Hc = H + d.h;
    const int c = cw + 4 * t;
    if (c + 4 <= d.lh) {
        uint4 o;
        o.x = plain_reduce(acc[0], m);
        o.y = plain_reduce(acc[1], m);
        o.z = plain_reduce(acc[2], m);
        o.w = plain_reduce(acc[3], m);
        if (((uintptr_t)(Hc + c) & 15) == 0) {
            *(uint4 *)(Hc + c) = o;
        } else {
            Hc[c] = o.x; Hc[c + 1] = o.y; Hc[c + 2] = o.z; Hc[c + 3] = o.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (c + r < d.lh) Hc[c + r] = plain_reduce(acc[r], m);
    }
}

// One tile of products of 2^K points (16384 >> K of them).
template <int K, class MD>
CM_DEV void task_tile(u32 *lds, const u32 *F, const u32 *G, u32 *H,
                                       const MulDesc *d, int cnt, const u32 *lv, const u32 *lvi,
                                       MD md, u32 sc) {
    mul_tile<K, kPW, kPT>(lds, lds + kPLW, F, G, H, d, cnt, lv, lvi, md, sc);
}

template <class MD>
__global__ __launch_bounds__(kPT) void k_list_persistent(
    const u32 *__restrict__ F, const u32 *__restrict__ G, u32 *__restrict__ H,
    const MulDesc *__restrict__ descs, const ListTask *__restrict__ tasks, int ntasks,
    int *__restrict__ counter, const u32 *__restrict__ scales, u32 *__restrict__ big_s,
    const u32 *__restrict__ lv, const u32 *__restrict__ lvi, MD md, PlainMod pm) {
    __shared__ __align__(16) u32 lds[2 * kPLW];
    __shared__ int s_next;
    if (threadIdx.x == 0) s_next = atomicAdd(counter, 1);
    __syncthreads();
    // uniform: the task and its descriptor load through the scalar unit
    const int ti = __builtin_amdgcn_readfirstlane(s_next);
    if (ti >= ntasks) return;   // one workgroup per task: never taken
    {
        const ListTask t = tasks[ti];
        if (t.kind == kTaskPlain) {
            task_plain(lds, F, G, H, descs + t.first, t.count, pm);
        } else if (t.kind == kTaskBigCyc) {
            task_big_half<false>(lds, F, G, H, descs + t.first, big_s, t.count, lv, lvi, md,
                                 scales[15]);
        } else if (t.kind == kTaskBigNeg) {
            task_big_half<true>(lds, F, G, H, descs + t.first, big_s, t.count, lv, lvi, md,
                                scales[15]);
        } else {
            const MulDesc *d = descs + t.first;
            const u32 sc = scales[t.logn];
#define CM_PMIX(K)                                                          \
    case K:                                                                 \
        task_tile<K>(lds, F, G, H, d, t.count, lv, lvi, md, sc);            \
        break;
            switch (t.logn) {
                CM_PMIX(4) CM_PMIX(5) CM_PMIX(6) CM_PMIX(7) CM_PMIX(8) CM_PMIX(9) CM_PMIX(10)
                CM_PMIX(11) CM_PMIX(12) CM_PMIX(13) CM_PMIX(14)
            default: break;
            }
#undef CM_PMIX
        }
    }
}

namespace {

// pointwise constants of the 2^e-point products, e = 0 .. 15 (those of the
// last modulus are kept; callers hold the API lock)
Scales scale_table(const Family *fam) {
    static u32 last_p = 0;
    static Scales t;
    if (last_p != fam->p) {
        for (int e = 0; e < 16; ++e) t.v[e] = mul_scale(e, fam->p);
        last_p = fam->p;
    }
    return t;
}

}  // namespace

bool set_list_tile_words(int words) {
    for (int v = 0; v < kMixVariants; ++v)
        if (kMixWordsOf[v] == words) {
            g_mix_variant = v;
            return true;
        }
    return false;
}

int list_tile_words() { return kMixWordsOf[g_mix_variant]; }

// Products whose shorter operand has at most this many coefficients run on
// the schoolbook kernel (plain.hip) in list multiplies: there the la * lb
// products cost less than three transforms of 2^ceil(log2(la + lb - 1))
// points.  0 sends everything to the NTT.  The tuner sets it.  Measured on
// the 4096 uneven pairs: 32 in the one-launch form (its transform tasks run
// at about two thirds of the VALU peak, its schoolbook tasks well below:
// 0.29 ms per call against 0.34 ms at 256), 256 in the per-class form.
int g_plain_cutoff = 32;

bool set_list_plain_cutoff(int c) {
    if (c < 0 || c > 1024) return false;
    g_plain_cutoff = c;
    return true;
}

int list_plain_cutoff() { return g_plain_cutoff; }

// The strided multiplies of tree and evaluation levels (equal shapes) take
// the schoolbook kernel up to this shorter-operand length: the measured
// crossover of equal-length pairs (plain faster up to 256 coefficients).
int strided_plain_cutoff() { return 256; }

// smallest class of 2^14-point products that runs in LDS (one workgroup per
// CU at a time, so it needs a few per CU to pay)
constexpr int64_t kLds14MinCount = 512;

// mixed-size tiles of mv = 0, 1, 2 (2048, 4096, 8192 words)
template <class MD>
static hipError_t launch_mixed(int mv, dim3 g, const u32 *F, const u32 *G, u32 *H,
                               const MulDesc *d_descs, const TileDesc *d_tiles, const Family *fam,
                               MD md, const Scales &sc, hipStream_t s) {
    if (mv == 0)
        hipLaunchKernelGGL((k_mul_mixed<2048, 128, MD>), g, dim3(128), 0, s, F, G, H, d_descs,
                           d_tiles, fam->lv, fam->lvi, md, sc);
    else if (mv == 1)
        hipLaunchKernelGGL((k_mul_mixed<4096, 256, MD>), g, dim3(256), 0, s, F, G, H, d_descs,
                           d_tiles, fam->lv, fam->lvi, md, sc);
    else
        hipLaunchKernelGGL((k_mul_mixed<8192, 512, MD>), g, dim3(512), 0, s, F, G, H, d_descs,
                           d_tiles, fam->lv, fam->lvi, md, sc);
    return hipGetLastError();
}

static hipError_t launch_list_mul_ntt(const MulDesc *descs, int64_t count, const u32 *F,
                                      const u32 *G, u32 *H, const Family *fam, hipStream_t s);
static hipError_t launch_list_persistent(const MulDesc *descs, int64_t count, const u32 *F,
                                         const u32 *G, u32 *H, const Family *fam, hipStream_t s);

// 1: the persistent launch (default); 0: one launch per size class
int g_list_persistent = [] {
    const char *e = getenv("CUMODP_LIST_PERSISTENT");
    return (e && *e == '0') ? 0 : 1;
}();

bool set_list_persistent(int on) {
    if (on != 0 && on != 1) return false;
    g_list_persistent = on;
    return true;
}

int list_persistent() { return g_list_persistent; }

hipError_t launch_list_mul(const MulDesc *descs, int64_t count, const u32 *F, const u32 *G,
                           u32 *H, const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (g_list_persistent) return launch_list_persistent(descs, count, F, G, H, fam, s);
    if (g_plain_cutoff <= 0) return launch_list_mul_ntt(descs, count, F, G, H, fam, s);
    std::vector<MulDesc> plain, rest;
    for (int64_t i = 0; i < count; ++i) {
        const MulDesc &d = descs[i];
        if (std::min(d.lf, d.lg) <= g_plain_cutoff) {
            MulDesc q = d;
            q.logn = d.lg > d.lf ? 1 : 0;   // sweep the shorter operand
            plain.push_back(q);
        } else {
            rest.push_back(d);
        }
    }
    if (!plain.empty()) {
        hipError_t e = launch_plain_list(plain.data(), (int64_t)plain.size(), F, G, H, fam->p, s);
        if (e != hipSuccess) return e;
    }
    return launch_list_mul_ntt(rest.data(), (int64_t)rest.size(), F, G, H, fam, s);
}

static hipError_t run_four_step_classes(const MulDesc *sorted, int64_t count,
                                        const MulDesc *d_descs, const u32 *F, const u32 *G,
                                        u32 *H, const Family *fam, hipStream_t s);

static hipError_t launch_list_mul_ntt(const MulDesc *descs, int64_t count, const u32 *F,
                                      const u32 *G, u32 *H, const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    // order: small sizes (tiled), then 2^13, then four-step classes by size
    // (a stable counting sort on logn)
    std::vector<MulDesc> sorted((size_t)count);
    {
        int64_t pos[34] = {0};
        for (int64_t i = 0; i < count; ++i) ++pos[descs[i].logn + 1];
        for (int c = 1; c < 34; ++c) pos[c] += pos[c - 1];
        for (int64_t i = 0; i < count; ++i) sorted[(size_t)pos[descs[i].logn]++] = descs[i];
    }
    // The tuned tile suits large batches.  A batch whose small products fill
    // fewer than two tiles per CU gets smaller tiles (more workgroups), down
    // to 2048 words; tile sizes below the largest small product are skipped.
    int mv = g_mix_variant;
    {
        int64_t small_words = 0;
        int max_small = 0;
        for (int64_t q = 0; q < count && sorted[(size_t)q].logn <= 13; ++q) {
            small_words += (int64_t)1 << sorted[(size_t)q].logn;
            max_small = std::max(max_small, (int)sorted[(size_t)q].logn);
        }
        while (mv > 0 && small_words / kMixWordsOf[mv] < 2 * 256 &&
               (1 << max_small) <= kMixWordsOf[mv - 1])
            --mv;
    }
    const int mix_words = kMixWordsOf[mv];
    std::vector<TileDesc> tiles;
    int64_t i = 0;
    while (i < count && ((int64_t)1 << sorted[(size_t)i].logn) <= mix_words) {
        const int lg = sorted[(size_t)i].logn;
        const int per = mix_words >> lg;
        int64_t j = i;
        while (j < count && sorted[(size_t)j].logn == lg && j - i < per) ++j;
        tiles.push_back(TileDesc{(int32_t)i, (int32_t)(j - i), lg, 0});
        i = j;
    }
    const int64_t n_small_end = i;
    int64_t n13_end = n_small_end;   // end of the one-product-per-workgroup sizes
    while (n13_end < count && sorted[(size_t)n13_end].logn <= kMaxLdsLog) ++n13_end;
    // a class of 2^14-point products large enough to fill the chip runs one
    // product per 1024-thread workgroup too (F and G tiles take 128 KiB of
    // LDS): no gather/scatter slabs and one launch instead of three
    {
        int64_t e14 = n13_end;
        while (e14 < count && sorted[(size_t)e14].logn == kMaxLdsLog + 1) ++e14;
        if (e14 - n13_end >= kLds14MinCount) n13_end = e14;
    }

    const size_t desc_bytes = sizeof(MulDesc) * (size_t)count;
    const size_t tile_bytes = sizeof(TileDesc) * tiles.size();
    std::vector<char> blob(desc_bytes + tile_bytes);
    memcpy(blob.data(), sorted.data(), desc_bytes);
    if (tile_bytes) memcpy(blob.data() + desc_bytes, tiles.data(), tile_bytes);
    char *dblob = (char *)workspace(7, blob.size(), s);
    if (!dblob) return hipErrorOutOfMemory;
    hipError_t e = stage_upload(blob.data(), blob.size(), dblob, s);
    if (e != hipSuccess) return e;
    const MulDesc *d_descs = (const MulDesc *)dblob;
    const TileDesc *d_tiles = (const TileDesc *)(dblob + desc_bytes);

    if (!tiles.empty()) {
        ProfScope ps(kProfMulMixed, s);
        const Scales sc = scale_table(fam);
        const dim3 g((unsigned)tiles.size());
        // values in [0, 2p) between butterflies when p < 2^30
        if (fam->p < (1u << 30))
            e = launch_mixed(mv, g, F, G, H, d_descs, d_tiles, fam, ModLazy{fam->p, fam->np}, sc, s);
        else
            e = launch_mixed(mv, g, F, G, H, d_descs, d_tiles, fam, Mod{fam->p, fam->np}, sc, s);
        if (e != hipSuccess) return e;
    }
    for (int64_t a = n_small_end; a < n13_end;) {
        const int lg = sorted[(size_t)a].logn;
        int64_t b = a;
        while (b < n13_end && sorted[(size_t)b].logn == lg) ++b;
        e = launch_mul_lds(lg, F, G, H, d_descs + a, (int)(b - a), fam, s);
        if (e != hipSuccess) return e;
        a = b;
    }
    return run_four_step_classes(sorted.data() + n13_end, count - n13_end, d_descs + n13_end, F,
                                 G, H, fam, s);
}

// The products sorted[0 .. count) (ascending logn, every logn > 13; their
// descriptors also at d_descs[0 .. count)) through batched four-step
// multiplies: per size, operands gathered into zero-padded slabs and products
// scattered back.
static hipError_t run_four_step_classes(const MulDesc *sorted, int64_t count,
                                        const MulDesc *d_descs, const u32 *F, const u32 *G,
                                        u32 *H, const Family *fam, hipStream_t s) {
    hipError_t e = hipSuccess;
    for (int64_t a = 0; a < count;) {
        const int lg = sorted[(size_t)a].logn;
        int64_t b = a;
        while (b < count && sorted[(size_t)b].logn == lg) ++b;
        const int64_t n = (int64_t)1 << lg;
        // process the class in chunks that keep the slabs within 1 GiB each
        const int64_t chunk = std::max<int64_t>(1, ((int64_t)1 << 28) / n);
        for (int64_t c0 = a; c0 < b; c0 += chunk) {
            const int cnt = (int)std::min<int64_t>(chunk, b - c0);
            u32 *sf = (u32 *)workspace(5, sizeof(u32) * (size_t)(n * cnt), s);
            u32 *sg = (u32 *)workspace(6, sizeof(u32) * (size_t)(n * cnt), s);
            if (!sf || !sg) return hipErrorOutOfMemory;
            // the slabs hold only the longest operand of each side (the
            // column pass zero-pads the rest, and skips the zero upper half
            // when both fit in n/2); the inverse pass stores the longest output
            int64_t mf = 1, mg = 1, mh = 1;
            for (int64_t q = c0; q < c0 + cnt; ++q) {
                mf = std::max<int64_t>(mf, sorted[(size_t)q].lf);
                mg = std::max<int64_t>(mg, sorted[(size_t)q].lg);
                mh = std::max<int64_t>(mh, sorted[(size_t)q].lh);
            }
            const unsigned gx = (unsigned)std::min<int64_t>(n / 1024, 64);
            {
            ProfScope ps(kProfGatherScatter, s);
            hipLaunchKernelGGL(k_gather, dim3(gx, cnt), dim3(1024), 0, s, sf, n, mf, F,
                               d_descs + c0, 0);
            hipLaunchKernelGGL(k_gather, dim3(gx, cnt), dim3(1024), 0, s, sg, n, mg, G,
                               d_descs + c0, 1);
            }
            e = hipGetLastError();
            FourStepIO io{sf, sg, mf, mg, n, sf, sg, sf, mh, n, cnt};
            if (e == hipSuccess) e = launch_mul_four_step(lg, io, fam, s);
            if (e != hipSuccess) return e;
            {
            ProfScope ps(kProfGatherScatter, s);
            hipLaunchKernelGGL(k_scatter, dim3(gx, cnt), dim3(1024), 0, s, H, sf, n,
                               d_descs + c0);
            }
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        a = b;
    }
    return hipSuccess;
}

// CUs of the current device: the persistent grid holds one workgroup per CU
// (its LDS allows no more).
static int device_cus() {
    static int cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            c <= 0)
            c = 256;
        cached[dev] = c;
    }
    return cached[dev];
}

// floor(log2(x)) of a cost, 0 .. 63: the exponent field of the double
// (a library ilogb call per task cost a few microseconds per list call)
static int cost_bucket(double x) {
    if (!(x >= 1.0)) return 0;
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    const int b = (int)((u >> 52) & 0x7ff) - 1023;
    return b > 63 ? 63 : b;
}

static hipError_t launch_list_persistent(const MulDesc *descs, int64_t count, const u32 *F,
                                         const u32 *G, u32 *H, const Family *fam, hipStream_t s) {
    if (count > INT32_MAX / 4) return hipErrorInvalidValue;
    constexpr int kBigLog = 15;
    const int cut = g_plain_cutoff;
    auto is_plain = [&](const MulDesc &d) { return cut > 0 && std::min(d.lf, d.lg) <= cut; };
    // descriptor order: the schoolbook products, then the transform sizes
    // ascending (class 0: schoolbook, class e: 2^e points)
    int64_t ncls[33] = {0};
    int64_t nplain_tasks = 0;
    for (int64_t i = 0; i < count; ++i) {
        const MulDesc &d = descs[i];
        if (is_plain(d)) {
            ++ncls[0];
            nplain_tasks += (d.lh + kPlainTaskOut - 1) / kPlainTaskOut;
        } else {
            ++ncls[d.logn];
        }
    }
    int64_t off[34];
    off[0] = 0;
    for (int c = 0; c < 33; ++c) off[c + 1] = off[c] + ncls[c];
    int64_t ntile = 0;
    for (int e = 4; e < kBigLog; ++e) ntile += (ncls[e] + (kPW >> e) - 1) / (kPW >> e);
    const int64_t nbig = ncls[kBigLog];
    const int64_t ntasks = nplain_tasks + ntile + 2 * nbig;
    if (ntasks > INT32_MAX) return hipErrorInvalidValue;

    // blob: the task counter (16 bytes), the pointwise scales of sizes 2^0 ..
    // 2^15 (64 bytes), the sorted descriptors, the tasks; written straight
    // into the pinned staging buffer
    const size_t scale_off = 16, desc_off = 80;
    const size_t task_off = desc_off + sizeof(MulDesc) * (size_t)count;
    const size_t bytes = task_off + sizeof(ListTask) * (size_t)ntasks;
    const Scales scl = scale_table(fam);
    std::vector<MulDesc> tail;   // host copy of the four-step classes' descriptors
    char *dblob = (char *)workspace(7, bytes, s);
    if (!dblob) return hipErrorOutOfMemory;
    hipError_t e = stage_upload_fill(bytes, dblob, s, [&](char *blob) {
        memset(blob, 0, desc_off);   // the task counter starts at 0
        memcpy(blob + scale_off, scl.v, sizeof(scl.v));
        MulDesc *sorted = (MulDesc *)(blob + desc_off);
        ListTask *tasks = (ListTask *)(blob + task_off);
        int64_t pos[33];
        for (int c = 0; c < 33; ++c) pos[c] = off[c];
        for (int64_t i = 0; i < count; ++i) {
            MulDesc q = descs[i];
            if (is_plain(q)) {
                q.logn = q.lg > q.lf ? 1 : 0;   // sweep the shorter operand
                sorted[pos[0]++] = q;
            } else {
                sorted[pos[q.logn]++] = q;
            }
        }
        // the tasks with their costs (butterflies; a schoolbook term costs
        // about an eighth of one), ordered by descending power-of-two cost
        // bucket: one pass counts the buckets, the second places the tasks
        auto each_task = [&](auto &&emit) {
            for (int64_t i = off[0]; i < off[1]; ++i) {
                const MulDesc &d = sorted[i];
                const int lb = std::min(d.lf, d.lg);
                for (int c0 = 0; c0 < d.lh; c0 += kPlainTaskOut) {
                    const int outs = std::min(kPlainTaskOut, d.lh - c0);
                    emit(ListTask{kTaskPlain, (int32_t)i, c0, 0}, (double)outs * lb / 8.0);
                }
            }
            for (int e2 = 4; e2 < kBigLog; ++e2) {
                const int per = kPW >> e2;
                for (int64_t i = off[e2]; i < off[e2 + 1]; i += per) {
                    const int cnt = (int)std::min<int64_t>(per, off[e2 + 1] - i);
                    emit(ListTask{kTaskTile, (int32_t)i, cnt, e2},
                         3.0 * e2 * (double)cnt * (1 << (e2 - 1)));
                }
            }
            for (int64_t i = off[kBigLog]; i < off[kBigLog + 1]; ++i) {   // two 2^14-point halves
                const int32_t slot = (int32_t)(i - off[kBigLog]);
                const double half = 3.0 * (kBigLog - 1) * (1 << (kBigLog - 2));
                emit(ListTask{kTaskBigCyc, (int32_t)i, slot, kBigLog}, half);
                emit(ListTask{kTaskBigNeg, (int32_t)i, slot, kBigLog}, half);
            }
        };
        int64_t nb[64] = {0};
        each_task([&](const ListTask &, double c) { ++nb[cost_bucket(c)]; });
        int64_t at[64];
        int64_t acc = 0;
        for (int b = 63; b >= 0; --b) {
            at[b] = acc;
            acc += nb[b];
        }
        each_task([&](const ListTask &t, double c) { tasks[at[cost_bucket(c)]++] = t; });
        tail.assign(sorted + off[kBigLog + 1], sorted + count);
    });
    if (e != hipSuccess) return e;
    int *d_counter = (int *)dblob;
    const u32 *d_scales = (const u32 *)(dblob + scale_off);
    const MulDesc *d_descs = (const MulDesc *)(dblob + desc_off);
    const ListTask *d_tasks = (const ListTask *)(dblob + task_off);
    u32 *big_s = nullptr;   // psi^-i B'/2 of each 2^15-point product, 2^14 words each
    if (nbig > 0) {
        big_s = (u32 *)workspace(20, sizeof(u32) * (size_t)(nbig << 14), s);
        if (!big_s) return hipErrorOutOfMemory;
    }
    if (ntasks > 0) {
        ProfScope ps(kProfListPersistent, s);
        const dim3 grid((unsigned)ntasks);
        const PlainMod pm = plain_mod(fam->p);
        if (fam->p < (1u << 30))
            hipLaunchKernelGGL(k_list_persistent<ModLazy>, grid, dim3(kPT), 0, s, F, G, H, d_descs,
                               d_tasks, (int)ntasks, d_counter, d_scales, big_s, fam->lv,
                               fam->lvi, ModLazy{fam->p, fam->np}, pm);
        else
            hipLaunchKernelGGL(k_list_persistent<Mod>, grid, dim3(kPT), 0, s, F, G, H, d_descs,
                               d_tasks, (int)ntasks, d_counter, d_scales, big_s, fam->lv,
                               fam->lvi, Mod{fam->p, fam->np}, pm);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nbig > 0) {
        ProfScope ps(kProfListPersistent, s);
        const dim3 grid(16, (unsigned)nbig);
        if (fam->p < (1u << 30))
            hipLaunchKernelGGL(k_big_combine<ModLazy>, grid, dim3(1024), 0, s, H,
                               d_descs + off[kBigLog], big_s, ModLazy{fam->p, fam->np});
        else
            hipLaunchKernelGGL(k_big_combine<Mod>, grid, dim3(1024), 0, s, H,
                               d_descs + off[kBigLog], big_s, Mod{fam->p, fam->np});
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // 2^16 points and more: batched four-step multiplies
    return run_four_step_classes(tail.data(), (int64_t)tail.size(), d_descs + off[kBigLog + 1], F,
                                 G, H, fam, s);
}

}  // namespace cm
