// batch.hip -- one call multiplying many pairs of uneven degree mod p.
//
// The reference has no list multiply for uneven degrees: list_stockham_dev
// (src/list_stockham.cu:595-630) transforms m rows of one length, and the
// unbuilt draft src/.list_fft_poly_mul.cu:15-53 multiplies equal-length
// pairs.  Here products are binned by transform size 2^e:
//   * e <= 12: one fused launch for every such product.  Each 256-thread
//     workgroup takes a 4096-word tile holding 4096 / 2^e products of one
//     size and runs the whole multiply in LDS; the tile's size selects the
//     code path at run time, so one launch covers all sizes;
//   * e == 13: one launch of the 8192-point workgroup multiply;
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

constexpr int kMixT = 256;          // threads of the mixed-size workgroup
constexpr int kMixWords = 4096;     // words per tile

struct Scales {
    u32 v[16];
};

template <int LOGL>
using MixTile = Tile<LOGL, (kMixWords >> LOGL), false, kMixT>;

template <int LOGL>
CM_DEV void mul_tile(u32 *A, u32 *B, const u32 *F, const u32 *G, u32 *H, const MulDesc *d,
                     int cnt, const u32 *lv, const u32 *lvi, Mod md, u32 sc) {
    using TL = MixTile<LOGL>;
    constexpr int L = TL::L;
    // element idx of the tile: vector c = idx >> LOGL, position i; all loads
    // of a thread are issued before the first LDS store
    constexpr int PER = TL::WORDS / kMixT;
    u32 fa[PER], ga[PER];
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
        if (c < cnt && i < d[c].lh) H[d[c].h + i] = A[TL::addr(c, i)];
    }
}

__global__ __launch_bounds__(kMixT) void k_mul_mixed(
    const u32 *__restrict__ F, const u32 *__restrict__ G, u32 *__restrict__ H,
    const MulDesc *__restrict__ descs, const TileDesc *__restrict__ tiles,
    const u32 *__restrict__ lv, const u32 *__restrict__ lvi, Mod md, Scales scales) {
    constexpr int W = kMixWords + (kMixWords >> 4);
    __shared__ u32 lds[2 * W];
    u32 *A = lds, *B = lds + W;
    const TileDesc t = tiles[blockIdx.x];
    const MulDesc *d = descs + t.first;
    const u32 sc = scales.v[t.logn];
    switch (t.logn) {
    case 4: mul_tile<4>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 5: mul_tile<5>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 6: mul_tile<6>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 7: mul_tile<7>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 8: mul_tile<8>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 9: mul_tile<9>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 10: mul_tile<10>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 11: mul_tile<11>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    case 12: mul_tile<12>(A, B, F, G, H, d, t.count, lv, lvi, md, sc); break;
    default: break;
    }
}

// zero-padded copy of operand `which` (0: F, 1: G) of descriptor row y into
// row y of a slab of n-word rows
__global__ void k_gather(u32 *__restrict__ slab, int64_t n, const u32 *__restrict__ src,
                         const MulDesc *__restrict__ d, int which) {
    const MulDesc dd = d[blockIdx.y];
    const int64_t off = which ? dd.g : dd.f;
    const int len = which ? dd.lg : dd.lf;
    u32 *row = slab + (int64_t)blockIdx.y * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
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

// Pinned staging for descriptors; reused once the previous upload finished.
struct Staging {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};

Staging &staging() {
    static Staging s;
    return s;
}

hipError_t stage_upload(const void *src, size_t bytes, void *dst, hipStream_t s) {
    Staging &st = staging();
    hipError_t e;
    if (!st.done) {
        e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    } else {
        e = hipEventSynchronize(st.done);
        if (e != hipSuccess) return e;
    }
    if (st.cap < bytes) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.cap = 0;
        e = hipHostMalloc(&st.host, bytes);
        if (e != hipSuccess) return e;
        st.cap = bytes;
    }
    memcpy(st.host, src, bytes);
    e = hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipEventRecord(st.done, s);
}

// pointwise constants of the 2^e-point products, e = 0 .. 15
Scales scale_table(const Family *fam) {
    Scales t;
    for (int e = 0; e < 16; ++e) t.v[e] = mul_scale(e, fam->p);
    return t;
}

}  // namespace

hipError_t launch_list_mul(const MulDesc *descs, int64_t count, const u32 *F, const u32 *G,
                           u32 *H, const Family *fam, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    // order: small sizes (tiled), then 2^13, then four-step classes by size
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return descs[a].logn < descs[b].logn; });
    std::vector<MulDesc> sorted((size_t)count);
    for (int64_t i = 0; i < count; ++i) sorted[(size_t)i] = descs[order[(size_t)i]];
    std::vector<TileDesc> tiles;
    int64_t i = 0;
    while (i < count && sorted[(size_t)i].logn <= 12) {
        const int lg = sorted[(size_t)i].logn;
        const int per = kMixWords >> lg;
        int64_t j = i;
        while (j < count && sorted[(size_t)j].logn == lg && j - i < per) ++j;
        tiles.push_back(TileDesc{(int32_t)i, (int32_t)(j - i), lg, 0});
        i = j;
    }
    const int64_t n_small_end = i;
    int64_t n13_end = n_small_end;
    while (n13_end < count && sorted[(size_t)n13_end].logn == 13) ++n13_end;

    const size_t desc_bytes = sizeof(MulDesc) * (size_t)count;
    const size_t tile_bytes = sizeof(TileDesc) * tiles.size();
    std::vector<char> blob(desc_bytes + tile_bytes);
    memcpy(blob.data(), sorted.data(), desc_bytes);
    if (tile_bytes) memcpy(blob.data() + desc_bytes, tiles.data(), tile_bytes);
    char *dblob = (char *)workspace(7, blob.size());
    if (!dblob) return hipErrorOutOfMemory;
    hipError_t e = stage_upload(blob.data(), blob.size(), dblob, s);
    if (e != hipSuccess) return e;
    const MulDesc *d_descs = (const MulDesc *)dblob;
    const TileDesc *d_tiles = (const TileDesc *)(dblob + desc_bytes);

    if (!tiles.empty()) {
        ProfScope ps(kProfMulMixed, s);
        hipLaunchKernelGGL(k_mul_mixed, dim3((unsigned)tiles.size()), dim3(kMixT), 0, s, F, G, H,
                           d_descs, d_tiles, fam->lv, fam->lvi, Mod{fam->p, fam->np},
                           scale_table(fam));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n13_end > n_small_end) {
        e = launch_mul_lds(13, F, G, H, d_descs + n_small_end, (int)(n13_end - n_small_end), fam,
                           s);
        if (e != hipSuccess) return e;
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
            u32 *sf = (u32 *)workspace(5, sizeof(u32) * (size_t)(n * cnt));
            u32 *sg = (u32 *)workspace(6, sizeof(u32) * (size_t)(n * cnt));
            if (!sf || !sg) return hipErrorOutOfMemory;
            const unsigned gx = (unsigned)std::min<int64_t>(n / 1024, 64);
            {
            ProfScope ps(kProfGatherScatter, s);
            hipLaunchKernelGGL(k_gather, dim3(gx, cnt), dim3(1024), 0, s, sf, n, F,
                               d_descs + c0, 0);
            hipLaunchKernelGGL(k_gather, dim3(gx, cnt), dim3(1024), 0, s, sg, n, G,
                               d_descs + c0, 1);
            }
            e = hipGetLastError();
            FourStepIO io{sf, sg, n, n, n, sf, sg, sf, n, n, cnt};
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
