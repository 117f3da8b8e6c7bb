// tree.hip -- subproduct tree over 2^k points, level by level.
//
// Reference: onlySubproductTree / subProductTree
// (src/fastPolyEvaluation.cu:512-656, :1411-1434).  The reference keeps two
// half-trees (Ml, Mr), leaves (-x_i, 1) with the raw negative, schoolbook
// products up to level 8 and FFT products with an implicit leading 1 above,
// and never forms the root.  Here one tree holds every level 0..k, each node
// the monic product of 2^i consecutive (x - x_j) stored with its leading
// coefficient (2^i + 1 canonical words), level i starting at word
//     off(i) = sum_{l < i} 2^(k-l) (2^l + 1).
// Level i comes from level i-1 by the same monic identity the reference uses
// (:601-650): with A = x^h + a and B = x^h + b (h = 2^(i-1), deg a, b < h),
//     A B = x^(2h) + x^h (a + b) + a b,
// where all products a b of a level are one call of the strided multiply:
// halves of up to 256 coefficients take the wavefront schoolbook kernel
// (plain.hip), larger ones the NTT multiply (one LDS workgroup per product,
// or four-step for 2^i > 2^13).  Levels 0 .. 8 are built before that in one
// launch, a 256-point subtree per workgroup (k_tree_bottom).
#include "accum.h"
#include "internal.h"

namespace cm {

int64_t tree_level_offset(int k, int i) {
    int64_t off = 0;
    for (int l = 0; l < i; ++l) off += ((int64_t)1 << (k - l)) * (((int64_t)1 << l) + 1);
    return off;
}

// after H[0 .. 2h-2] = a b: add (a + b) x^h and the leading 1
__global__ void k_tree_combine(u32 *__restrict__ tree, int64_t off_prev, int64_t off_cur, int h,
                               int64_t nprod, u32 p) {
    const int len = 2 * h + 1;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nprod * (h + 1);
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g / (h + 1);
        const int t = h + (int)(g % (h + 1));        // h .. 2h
        const u32 *A = tree + off_prev + (2 * j) * (h + 1);
        const u32 *B = A + (h + 1);
        u32 *C = tree + off_cur + j * len;
        if (t == 2 * h) {
            C[t] = 1u;
        } else {
            const u32 s = cm_add(A[t - h], B[t - h], p);
            C[t] = (t < 2 * h - 1) ? cm_add(C[t], s, p) : s;
        }
    }
}

// Levels 0 .. pin of the tree in one launch: workgroup J (blockDim.x =
// 2^pin threads, pin <= kTreeFuse) builds the subtree over points
// J 2^pin .. J 2^pin + 2^pin - 1 in LDS, level by level, and writes every
// level's nodes (a contiguous block of each level) to the tree.  Node
// coefficient r of level i (children a, b of level i-1, h + 1 words each,
// h = 2^(i-1)) is sum_u a_u b_(r-u), a schoolbook sum in a 64-bit
// accumulator (accum.h); the leading coefficients are 1 and come out of the
// same sum.
constexpr int kTreeFuse = 8;

__global__ __launch_bounds__(1 << kTreeFuse) void k_tree_bottom(const u32 *__restrict__ pts,
                                                               int pin, PlainMod pm,
                                                               u32 *__restrict__ tree,
                                                               int64_t n) {
    constexpr int BMAX = 1 << kTreeFuse;
    __shared__ u32 buf[2][2 * BMAX];
    const int B = 1 << pin;
    const int64_t J = blockIdx.x;
    const int t = threadIdx.x;
    {
        const u32 x = pts[J * B + t] % pm.p;
        const u32 lo = x ? pm.p - x : 0u;
        buf[0][2 * t] = lo;
        buf[0][2 * t + 1] = 1u;
        tree[2 * (J * B + t)] = lo;
        tree[2 * (J * B + t) + 1] = 1u;
    }
    int64_t off = 0;                          // word offset of level i - 1 in the tree
    for (int i = 1; i <= pin; ++i) {
        __syncthreads();
        const int h = 1 << (i - 1);
        const int len = 2 * h + 1, nodes = B >> i;
        off += (n >> (i - 1)) * (h + 1);      // level i
        const u32 *prev = buf[(i - 1) & 1];
        u32 *cur = buf[i & 1];
        u32 *out = tree + off + (J * nodes) * len;
        for (int x = t; x < nodes * len; x += B) {
            const int c = x / len, r = x - c * len;
            const u32 *a = prev + (2 * c) * (h + 1), *b = a + (h + 1);
            const int lo = r > h ? r - h : 0, hi = r < h ? r : h;
            const u32 v = plain_dot<1, -1>(a + lo, b + (r - lo), hi - lo + 1, pm);
            cur[x] = v;
            out[x] = v;
        }
    }
}

static int tblocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

hipError_t launch_subproduct_tree(const u32 *pts, int k, u32 p, u32 *tree, const Family *fam,
                                  hipStream_t s) {
    const int64_t n = (int64_t)1 << k;
    // levels 0 .. pin in one launch, one workgroup per 2^pin points
    const int pin = k < kTreeFuse ? k : kTreeFuse;
    hipLaunchKernelGGL(k_tree_bottom, dim3((unsigned)(n >> pin)), dim3(1 << pin), 0, s, pts, pin,
                       plain_mod(p), tree, n);
    hipError_t e = hipGetLastError();
    for (int i = pin + 1; i <= k && e == hipSuccess; ++i) {
        const int h = 1 << (i - 1);
        const int64_t nprod = n >> i;
        const int64_t op = tree_level_offset(k, i - 1), oc = tree_level_offset(k, i);
        // the level's products share one shape: a = node 2j and b = node 2j+1
        // of level i-1 (their non-leading h coefficients, 2(h+1) words apart),
        // a b (2h-1 words) at node j of level i; no descriptors.  Up to 256
        // coefficients each the strided multiply takes the schoolbook kernel
        // (faster than the transform there), above it the NTT multiply on
        // 2h points
        e = launch_mul_strided(i, tree + op, 2 * (h + 1), h, tree + op + (h + 1), 2 * (h + 1), h,
                               tree + oc, 2 * h + 1, 2 * h - 1, (int)nprod, fam, s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_tree_combine, dim3(tblocks(nprod * (h + 1))), dim3(256), 0, s, tree,
                           op, oc, h, nprod, p);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace cm
