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
#include <cstring>
#include <vector>

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

namespace {

// pointwise constants of the 2^e-point products, e = 0 .. 15
Scales scale_table(const Family *fam) {
    Scales t;
    for (int e = 0; e < 16; ++e) t.v[e] = mul_scale(e, fam->p);
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
// points.  0 sends everything to the NTT.  The tuner sets it.
int g_plain_cutoff = 256;

bool set_list_plain_cutoff(int c) {
    if (c < 0 || c > 1024) return false;
    g_plain_cutoff = c;
    return true;
}

int list_plain_cutoff() { return g_plain_cutoff; }

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

hipError_t launch_list_mul(const MulDesc *descs, int64_t count, const u32 *F, const u32 *G,
                           u32 *H, const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
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

static hipError_t launch_list_mul_ntt(const MulDesc *descs, int64_t count, const u32 *F,
                                      const u32 *G, u32 *H, const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    // order: small sizes (tiled), then 2^13, then four-step classes by size
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return descs[a].logn < descs[b].logn; });
    std::vector<MulDesc> sorted((size_t)count);
    for (int64_t i = 0; i < count; ++i) sorted[(size_t)i] = descs[order[(size_t)i]];
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
    for (int64_t a = n13_end; a < count;) {
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

}  // namespace cm
