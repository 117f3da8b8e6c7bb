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
// where all products a b of a level are one call of a batched multiply:
// levels up to 5 use a one-thread-per-coefficient schoolbook kernel, levels
// whose halves have up to 256 coefficients the wavefront schoolbook kernel
// (plain.hip), larger ones the NTT list multiply (mixed-size LDS kernel, or
// four-step for 2^i > 2^13).
#include <vector>

#include "internal.h"

namespace cm {

int64_t tree_level_offset(int k, int i) {
    int64_t off = 0;
    for (int l = 0; l < i; ++l) off += ((int64_t)1 << (k - l)) * (((int64_t)1 << l) + 1);
    return off;
}

__global__ void k_tree_leaves(const u32 *__restrict__ pts, int64_t n, u32 p,
                              u32 *__restrict__ tree) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const u32 x = pts[j] % p;
        tree[2 * j] = x ? p - x : 0u;
        tree[2 * j + 1] = 1u;
    }
}

// schoolbook level: node j of level i (length 2h+1) from nodes 2j, 2j+1 of
// level i-1 (length h+1); one thread per output coefficient
__global__ void k_tree_plain(u32 *__restrict__ tree, int64_t off_prev, int64_t off_cur, int h,
                             int64_t nprod, u32 p) {
    const int len = 2 * h + 1;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nprod * len;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g / len;
        const int t = (int)(g % len);
        const u32 *A = tree + off_prev + (2 * j) * (h + 1);
        const u32 *B = A + (h + 1);
        u64 acc = 0;
        const int lo = t > h ? t - h : 0, hi = t < h ? t : h;
        for (int u = lo; u <= hi; ++u) acc = (acc + (u64)A[u] * B[t - u]) % p;
        tree[off_cur + j * len + t] = (u32)acc;
    }
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

static int tblocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

hipError_t launch_subproduct_tree(const u32 *pts, int k, u32 p, u32 *tree, const Family *fam,
                                  hipStream_t s) {
    const int64_t n = (int64_t)1 << k;
    hipLaunchKernelGGL(k_tree_leaves, dim3(tblocks(n)), dim3(256), 0, s, pts, n, p, tree);
    hipError_t e = hipGetLastError();
    for (int i = 1; i <= k && e == hipSuccess; ++i) {
        const int h = 1 << (i - 1);
        const int64_t nprod = n >> i;
        const int64_t op = tree_level_offset(k, i - 1), oc = tree_level_offset(k, i);
        if (i <= 5) {
            hipLaunchKernelGGL(k_tree_plain, dim3(tblocks(nprod * (2 * h + 1))), dim3(256), 0, s,
                               tree, op, oc, h, nprod, p);
            e = hipGetLastError();
            continue;
        }
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
