// eval.hip -- fast multipoint evaluation of a polynomial at 2^k points.
//
// Reference: subInvTree and fastEvaluation (src/fastPolyEvaluation.cu:666-1400,
// driver :1371-1400).  The reference evaluates F of degree < 2^k at 2^k points
// by going down the subproduct tree, replacing each remainder by its
// remainders modulo the two children.  It uses Newton-iteration inverses of
// the reversed node polynomials (the "subinverse tree") for the divisions
// above level plainMulLimit and plain division below.  Here:
//   * the subproduct tree is built by launch_subproduct_tree (tree.hip);
//   * subinverse tree: I_i(j) = rev(M_i(j))^-1 mod x^(2^i).  At level L0 it
//     comes from a direct power-series inversion (one thread per node).
//     Above that, I_(i+1)(j) comes from the two children: their product is
//     the inverse mod x^(2^i), and one Newton step g <- g (2 - f g) doubles
//     the precision to 2^(i+1).  Only the high half of f g is needed (its
//     low half is 1), so that is three batched products on 2^(i+1) points
//     per level and no elementwise pass;
//   * remainder tree: R_k = F, and R_i(j) = R_(i+1)(j/2) mod M_i(j) by
//     q = rev(rev(A) I_i(j) mod x^h) and R = (A - q M) mod x^h with h = 2^i.
//     That is two batched products per level;
//   * at level L0 each remainder (2^L0 coefficients) is evaluated at its
//     node's 2^L0 points by Horner's rule.
// Every batched product of a level is one launch_mul_strided call: the
// level's products share one shape, so they need no descriptors (one LDS
// multiply per workgroup up to 2^13 points, batched four-step above).
#include "accum.h"
#include "internal.h"

namespace cm {

constexpr int kEvalBase = 4;   // L0: levels below use plain arithmetic

static int eblocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// I_L0(j) = rev(M)^-1 mod x^H for the nodes of level L0 (H = 2^L0): with
// f = rev(M) (f_0 = 1, f_s = M_(H-s)), g_0 = 1 and g_t = -sum_(s=1..t) f_s g_(t-s).
// One thread per node, node and inverse in registers; the sums run in 64-bit
// accumulators folded every four products (accum.h).
template <int H>
__global__ void k_inv_base(const u32 *__restrict__ M, int64_t nodes, PlainMod pm,
                           u32 *__restrict__ inv) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nodes;
         j += (int64_t)gridDim.x * blockDim.x) {
        const u32 *m = M + j * (H + 1);
        u32 f[H], g[H];
#pragma unroll
        for (int s = 1; s < H; ++s) f[s] = m[H - s];
        g[0] = 1u;
#pragma unroll
        for (int t = 1; t < H; ++t) {
            u64 acc = 0;
#pragma unroll
            for (int s = 1; s <= t; ++s) {
                acc += (u64)f[s] * g[t - s];
                if ((s & 3) == 0) acc = plain_fold(acc, pm.c32);
            }
            const u32 r = plain_reduce(acc, pm);
            g[t] = r ? pm.p - r : 0u;
        }
        u32 *o = inv + j * H;
#pragma unroll
        for (int t = 0; t < H; ++t) o[t] = g[t];
    }
}

// Negated reversed nodes of every level the subinverse tree needs, in one
// launch: for level l = first + blockIdx.y (node length h2 + 1, h2 = 2^l),
// f(j) = -rev(M_l(j)) mod x^h2, f_t = -M_(h2 - t) for t < h2, stored at
// out + blockIdx.y n (each level's reversed nodes take n words).
struct LevelOffsets {
    int64_t off[32];   // word offset of level l in the tree
};

__global__ void k_rev_nodes(const u32 *__restrict__ tree, LevelOffsets lo, int first, int64_t n,
                            u32 p, u32 *__restrict__ out) {
    const int l = first + (int)blockIdx.y;
    const int h2 = 1 << l;
    const u32 *M = tree + lo.off[l];
    u32 *f = out + (int64_t)blockIdx.y * n;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g >> l;
        const int t = (int)(g & (h2 - 1));
        const u32 v = M[j * (h2 + 1) + (h2 - t)];
        f[g] = v ? p - v : 0u;
    }
}

// reversed top half of each child's dividend: out(j)_t = A(j/2)_(2h-1-t), t < h
__global__ void k_div_rev_top(const u32 *__restrict__ A, int h, int64_t nodes,
                              u32 *__restrict__ out) {
    const int64_t total = nodes * h;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g >> __builtin_ctz((unsigned)h);    // h is a power of two
        const int t = (int)(g & (h - 1));
        out[g] = A[(j >> 1) * (2 * h) + (2 * h - 1 - t)];
    }
}

// q(j)_t = qrev(j)_(h-1-t)
__global__ void k_reverse(const u32 *__restrict__ in, int h, int64_t nodes,
                          u32 *__restrict__ out) {
    const int64_t total = nodes * h;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g >> __builtin_ctz((unsigned)h);
        const int t = (int)(g & (h - 1));
        out[g] = in[j * h + (h - 1 - t)];
    }
}

// R(j)_t = A(j/2)_t - qm(j)_t, t < h.  With next != nullptr, also the next
// level's reversed top halves (k_div_rev_top of R at h/2): both children c
// of node j get next(c)_t' = R(j)_(h-1-t') for t' < h/2.
__global__ void k_div_finish(const u32 *__restrict__ A, const u32 *__restrict__ qm, int h,
                             int64_t nodes, u32 p, u32 *__restrict__ R, u32 *__restrict__ next) {
    const int64_t total = nodes * h;
    const int hh = h >> 1;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = g >> __builtin_ctz((unsigned)h);
        const int t = (int)(g & (h - 1));
        const u32 r = cm_sub(A[(j >> 1) * (2 * h) + t], qm[g], p);
        R[g] = r;
        if (next && t >= hh) {
            next[(2 * j) * hh + (h - 1 - t)] = r;
            next[(2 * j + 1) * hh + (h - 1 - t)] = r;
        }
    }
}

// values_x = R_L0(x / h)(point x) by Horner's rule, one Montgomery product
// by z 2^32 per coefficient
__global__ void k_horner(const u32 *__restrict__ R, int h, const u32 *__restrict__ pts,
                         int64_t n, PlainMod pm, u32 *__restrict__ val) {
    const int lh = __builtin_ctz((unsigned)h);   // h is a power of two
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const u32 *r = R + ((x >> lh) << lh);
        const u32 zr = cm_mul_tw(pts[x] % pm.p, pm.r2, pm.p, pm.np);   // z 2^32 mod p
        u32 acc = 0;
        for (int t = h - 1; t >= 0; --t)
            acc = cm_add(cm_mul_tw(acc, zr, pm.p, pm.np), r[t], pm.p);
        val[x] = acc;
    }
}

// The lowest remainder levels and the leaves in one launch.  Workgroup J
// takes node J of level `pin` (its remainder at A + J 2^pin, blockDim.x =
// kEvalParts 2^pin threads, pin <= kEvalFuse) and descends its subtree to
// level L0 in LDS: at level i (h = 2^i) slot t = c h + u computes coefficient u of
// node c, first qrev_u = sum_(s<=u) A'_(2h-1-s) I_(u-s) (A' the parent's
// remainder, I = I_i(node)), then q_s = qrev_(h-1-s) and
// R_u = A'_u - sum_(s<=u) q_s M_(u-s) -- the same division as the level-wide
// launches, with the products as schoolbook sums in 64-bit accumulators
// (accum.h).  Then thread t evaluates the leaf holding point J 2^pin + t by
// Horner's rule, as k_horner does.
constexpr int kEvalFuse = 8;

// Each sum is split over kEvalParts threads (thread = part 2^pin + t) whose
// residues thread t adds up: 4 2^pin threads per workgroup.
constexpr int kEvalParts = 4;
static_assert(kEvalParts == 4, "the partial sums are added as four terms");

__global__ __launch_bounds__(kEvalParts << kEvalFuse) void k_rem_descend(
    const u32 *__restrict__ A, int pin, int L0, const u32 *__restrict__ tree, LevelOffsets lo,
    const u32 *__restrict__ inv, int64_t n, const u32 *__restrict__ pts, PlainMod pm,
    u32 *__restrict__ val) {
    constexpr int BMAX = 1 << kEvalFuse;
    __shared__ u32 cur[BMAX], qb[BMAX], Il[BMAX], Ml[BMAX + (BMAX >> 4)], pp[kEvalParts * BMAX];
    const int B = 1 << pin;
    const int64_t J = blockIdx.x;
    const int tid = threadIdx.x, t = tid & (B - 1), part = tid >> pin;
    if (part == 0) cur[t] = A[J * B + t];
    for (int i = pin - 1; i >= L0; --i) {
        const int h = 1 << i, nodes = B >> i;
        const int c = t >> i, u = t & (h - 1);
        // this level's inverses (B words) and node polynomials (nodes (h+1)
        // words) of the subtree, both contiguous in global memory
        if (part == 0) Il[t] = inv[(int64_t)(i - L0) * n + J * B + t];
        const u32 *Mg = tree + lo.off[i] + (J * nodes) * (h + 1);
        for (int x = tid; x < nodes * (h + 1); x += kEvalParts * B) Ml[x] = Mg[x];
        __syncthreads();
        const u32 *par = cur + (c >> 1) * (2 * h);
        const u32 *I = Il + c * h;
        pp[tid] = plain_dot_part<-1, -1, kEvalParts>(par + 2 * h - 1, I + u, u + 1, part, pm);
        __syncthreads();
        if (part == 0)
            qb[t] = cm_add(cm_add(pp[t], pp[B + t], pm.p), cm_add(pp[2 * B + t], pp[3 * B + t], pm.p),
                           pm.p);
        __syncthreads();
        const u32 *q = qb + c * h, *M = Ml + c * (h + 1);
        pp[tid] = plain_dot_part<-1, -1, kEvalParts>(q + h - 1, M + u, u + 1, part, pm);
        __syncthreads();
        u32 r = 0;
        if (part == 0) {
            const u32 qm = cm_add(cm_add(pp[t], pp[B + t], pm.p),
                                  cm_add(pp[2 * B + t], pp[3 * B + t], pm.p), pm.p);
            r = cm_sub(par[u], qm, pm.p);
        }
        __syncthreads();                      // every read of cur (par) is done
        if (part == 0) cur[t] = r;
        __syncthreads();
    }
    if (part != 0) return;
    const int64_t x = J * B + t;
    const u32 *r = cur + ((t >> L0) << L0);
    const u32 zr = cm_mul_tw(pts[x] % pm.p, pm.r2, pm.p, pm.np);   // z 2^32 mod p
    u32 acc = 0;
    for (int s2 = (1 << L0) - 1; s2 >= 0; --s2)
        acc = cm_add(cm_mul_tw(acc, zr, pm.p, pm.np), r[s2], pm.p);
    val[x] = acc;
}

// The lowest subinverse levels in one launch: workgroup J (blockDim.x =
// kEvalParts 2^pin threads, kEvalBase < pin <= kEvalFuse) takes the nodes of level
// kEvalBase .. pin under node J of level pin.  The base level comes from the
// power-series recurrence of k_inv_base (one thread per node); each level
// above from the Newton step of the level-wide launches, with the three
// products as schoolbook sums in LDS: slot t = c 2h + v (node c of level
// i+1, h = 2^i) forms g0_v (v < h), then -E_(v-h) = -sum f_s g0_(v-s)
// (v >= h, f_s = M_(2h-s)), then the high half sum g0_s (-E)_(v-h-s).
// Every level is written to the subinverse tree.
__global__ __launch_bounds__(kEvalParts << kEvalFuse) void k_inv_bottom(
    const u32 *__restrict__ tree, LevelOffsets lo, int pin, PlainMod pm, u32 *__restrict__ inv,
    int64_t n) {
    constexpr int BMAX = 1 << kEvalFuse, H0 = 1 << kEvalBase;
    __shared__ u32 cur[BMAX], nxt[BMAX], tmp[BMAX / 2], Ml[BMAX + BMAX / (2 * H0)],
        pp[kEvalParts * BMAX];
    const int B = 1 << pin;
    const int64_t J = blockIdx.x;
    const int tid = threadIdx.x, t = tid & (B - 1), part = tid >> pin;
    const u32 p = pm.p;
    auto sum4 = [&](int t2) {
        return cm_add(cm_add(pp[t2], pp[B + t2], p), cm_add(pp[2 * B + t2], pp[3 * B + t2], p), p);
    };
    const int nb = B >> kEvalBase;
    if (part == 0 && t < nb) {
        const u32 *m = tree + lo.off[kEvalBase] + (J * nb + t) * (H0 + 1);
        u32 f[H0], g[H0];
#pragma unroll
        for (int s2 = 1; s2 < H0; ++s2) f[s2] = m[H0 - s2];
        g[0] = 1u;
#pragma unroll
        for (int w = 1; w < H0; ++w) {
            u64 acc = 0;
#pragma unroll
            for (int s2 = 1; s2 <= w; ++s2) {
                acc += (u64)f[s2] * g[w - s2];
                if ((s2 & 3) == 0) acc = plain_fold(acc, pm.c32);
            }
            const u32 r = plain_reduce(acc, pm);
            g[w] = r ? p - r : 0u;
        }
#pragma unroll
        for (int w = 0; w < H0; ++w) cur[t * H0 + w] = g[w];
    }
    __syncthreads();
    if (part == 0) inv[J * B + t] = cur[t];
    for (int i = kEvalBase; i < pin; ++i) {
        const int h = 1 << i, nodes = B >> (i + 1);
        const int c = t >> (i + 1), v = t & (2 * h - 1);
        const u32 *Mg = tree + lo.off[i + 1] + (J * nodes) * (2 * h + 1);
        for (int x = tid; x < nodes * (2 * h + 1); x += kEvalParts * B) Ml[x] = Mg[x];
        u32 *g0 = nxt + c * (2 * h);
        if (v < h) {
            const u32 *ia = cur + (2 * c) * h, *ib = ia + h;
            pp[tid] = plain_dot_part<1, -1, kEvalParts>(ia, ib + v, v + 1, part, pm);
        }
        __syncthreads();
        if (part == 0 && v < h) g0[v] = sum4(t);
        __syncthreads();
        if (v >= h) {
            const u32 *M = Ml + c * (2 * h + 1);
            // s = v - h + 1 + k: M_(2h - s) g0_(v - s) = M_(3h-1-v-k) g0_(h-1-k)
            pp[tid] = plain_dot_part<-1, -1, kEvalParts>(M + (3 * h - 1 - v), g0 + (h - 1), h,
                                                         part, pm);
        }
        __syncthreads();
        if (part == 0 && v >= h) {
            const u32 e = sum4(t);
            tmp[c * h + (v - h)] = e ? p - e : 0u;
        }
        __syncthreads();
        if (v >= h) {
            const int w = v - h;
            pp[tid] = plain_dot_part<1, -1, kEvalParts>(g0, tmp + c * h + w, w + 1, part, pm);
        }
        __syncthreads();
        if (part == 0 && v >= h) g0[v] = sum4(t);   // high half; the sums read the low half
        __syncthreads();
        if (part == 0) {
            cur[t] = nxt[t];
            inv[(int64_t)(i + 1 - kEvalBase) * n + J * B + t] = nxt[t];
        }
        __syncthreads();
    }
}

// F mod p: the remainder tree reads this copy, so F may hold any 32-bit words
__global__ void k_reduce_copy(const u32 *__restrict__ F, int64_t n, u32 p, u32 *__restrict__ out) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x)
        out[g] = F[g] % p;
}

// R_i(j) = R_(i+1)(j/2) mod M_i(j) for the levels below L0 (H = 2^i <= 8),
// one thread per node, by schoolbook division by the monic M_i(j); used only
// when every remainder level is kept
template <int H>
__global__ void k_rem_small(const u32 *__restrict__ parent, const u32 *__restrict__ M,
                            int64_t nodes, u32 p, u32 *__restrict__ out) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nodes;
         j += (int64_t)gridDim.x * blockDim.x) {
        u32 r[2 * H];
#pragma unroll
        for (int t = 0; t < 2 * H; ++t) r[t] = parent[(j >> 1) * (2 * H) + t];
        const u32 *m = M + j * (H + 1);
#pragma unroll
        for (int d = 2 * H - 1; d >= H; --d) {
            const u64 c = r[d] ? p - r[d] : 0u;
#pragma unroll
            for (int t = 0; t < H; ++t)
                r[d - H + t] = (u32)((r[d - H + t] + c * m[t]) % p);
        }
#pragma unroll
        for (int t = 0; t < H; ++t) out[j * H + t] = r[t];
    }
}

int64_t eval_scratch_words(int k) {
    const int64_t n = (int64_t)1 << k;
    const int L0 = k < kEvalBase ? k : kEvalBase;
    const int nrev = k - 1 - L0 > 0 ? k - 1 - L0 : 0;          // reversed-node levels
    return tree_level_offset(k, k + 1) + (int64_t)(k - L0 + 1) * n + 6 * n + (int64_t)nrev * n;
}

int64_t eval_inverse_offset(int k) { return tree_level_offset(k, k + 1); }

// Level 0 in the reference's half-tree layout (leavesSubproductTree,
// src/fastPolyEvaluation.cu:511-518): node j of the left half is
// (-x_j, 1) and of the right half (-x_(nh+j), 1), -x the 32-bit negation
__global__ void k_ref_leaves(const u32 *__restrict__ pts, int64_t nh, u32 *__restrict__ left,
                             u32 *__restrict__ right) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh;
         j += (int64_t)gridDim.x * blockDim.x) {
        left[2 * j] = 0u - pts[j];
        left[2 * j + 1] = 1u;
        right[2 * j] = 0u - pts[nh + j];
        right[2 * j + 1] = 1u;
    }
}

hipError_t launch_ref_leaves(const u32 *pts, int64_t nh, u32 *left, u32 *right, hipStream_t s) {
    hipLaunchKernelGGL(k_ref_leaves, dim3(eblocks(nh)), dim3(256), 0, s, pts, nh, left, right);
    return hipGetLastError();
}

hipError_t launch_fast_evaluation(const u32 *F_in, const u32 *pts, int k, u32 p, u32 *values,
                                  u32 *scratch, const Family *fam, hipStream_t s,
                                  const EvalKeep *keep) {
    const int64_t n = (int64_t)1 << k;
    const int L0 = k < kEvalBase ? k : kEvalBase;
    u32 *tree = scratch;
    u32 *inv = tree + tree_level_offset(k, k + 1);           // levels L0 .. k, n words each
    u32 *R0 = inv + (int64_t)(k - L0 + 1) * n, *R1 = R0 + n;
    u32 *T1 = R1 + n, *T2 = T1 + n, *T3 = T2 + n;
    u32 *Fr = T3 + n;                                        // F mod p
    u32 *rev = Fr + n;                                       // levels L0+1 .. k-1, n words each
    auto inv_at = [&](int i) { return inv + (int64_t)(i - L0) * n; };
    auto node = [&](int i) { return tree + tree_level_offset(k, i); };

    hipError_t e = launch_subproduct_tree(pts, k, p, tree, fam, s);
    if (e != hipSuccess) return e;
    // ---- subinverse tree, levels L0 .. k-1: levels L0 .. pinv in one launch
    // (k_inv_bottom) when pinv = min(k - 1, kEvalFuse) > L0, else the base
    // level alone (k_inv_base); Newton steps up from there
    const int pinv = k - 1 < kEvalFuse ? k - 1 : kEvalFuse;
    int istart = L0;
    if (pinv > L0) {
        LevelOffsets lo{};
        for (int l = 0; l <= k && l < 32; ++l) lo.off[l] = tree_level_offset(k, l);
        hipLaunchKernelGGL(k_inv_bottom, dim3((unsigned)(n >> pinv)), dim3(kEvalParts << pinv), 0, s,
                           tree, lo, pinv, plain_mod(p), inv, n);
        e = hipGetLastError();
        istart = pinv;
    } else {
        const int h = 1 << L0;
        static_assert(kEvalBase == 4, "base inverse kernel is instantiated for 16 words");
        if (h == 16)
            hipLaunchKernelGGL(k_inv_base<16>, dim3(eblocks(n >> L0)), dim3(256), 0, s, node(L0),
                               n >> L0, plain_mod(p), inv_at(L0));
        else if (h == 8)
            hipLaunchKernelGGL(k_inv_base<8>, dim3(eblocks(n >> L0)), dim3(256), 0, s, node(L0),
                               n >> L0, plain_mod(p), inv_at(L0));
        else if (h == 4)
            hipLaunchKernelGGL(k_inv_base<4>, dim3(eblocks(n >> L0)), dim3(256), 0, s, node(L0),
                               n >> L0, plain_mod(p), inv_at(L0));
        else if (h == 2)
            hipLaunchKernelGGL(k_inv_base<2>, dim3(eblocks(n >> L0)), dim3(256), 0, s, node(L0),
                               n >> L0, plain_mod(p), inv_at(L0));
        else
            hipLaunchKernelGGL(k_inv_base<1>, dim3(eblocks(n >> L0)), dim3(256), 0, s, node(L0),
                               n >> L0, plain_mod(p), inv_at(L0));
        e = hipGetLastError();
    }
    if (e == hipSuccess && k - 1 > L0) {
        LevelOffsets lo{};
        for (int l = 0; l <= k && l < 32; ++l) lo.off[l] = tree_level_offset(k, l);
        hipLaunchKernelGGL(k_rev_nodes, dim3(eblocks(n), k - 1 - L0), dim3(256), 0, s, tree, lo,
                           L0 + 1, n, p, rev);
        e = hipGetLastError();
    }
    // Newton step from level i to i + 1 (h = 2^i): with f = rev(M_(i+1)(j))
    // and g0 = its inverse mod x^h, f g0 = 1 + x^h E mod x^(2h), and
    // g0 (2 - f g0) = g0 - x^h (g0 E) mod x^(2h).  So I_(i+1)(j) keeps g0 as
    // its low half and takes -(g0 E) mod x^h as its high half: three
    // products on 2h points per level, no elementwise pass.
    for (int i = istart; i + 1 <= k - 1 && e == hipSuccess; ++i) {
        const int h = 1 << i;
        const int N = (int)(n >> (i + 1));                   // nodes of level i + 1
        u32 *I1 = inv_at(i + 1);                             // node j at I1 + 2h j
        // g0(j) = I_i(2j) I_i(2j+1) mod x^h (rev is multiplicative), the
        // low half of I_(i+1)(j)
        e = launch_mul_strided(i + 1, inv_at(i), 2 * h, h, inv_at(i) + h, 2 * h, h, I1, 2 * h, h,
                               N, fam, s);
        if (e != hipSuccess) break;
        // (-f) g0: coefficients h .. 2h-1 are -E.  The cyclic product on 2h
        // points gives them exactly (f g0 has degree < 3h - 1, so the wrapped
        // terms land below x^(h-1)); short products take the schoolbook
        // kernel on 4h points instead (same words h .. 2h-1)
        const u32 *fneg = rev + (int64_t)(i - L0) * n;
        const int ln2 = h <= strided_plain_cutoff() ? i + 2 : i + 1;
        e = launch_mul_strided(ln2, fneg, 2 * h, 2 * h, I1, 2 * h, h, T3, 2 * h, 2 * h, N, fam, s);
        if (e != hipSuccess) break;
        // high half of I_(i+1)(j) = g0 (-E) mod x^h (reads the low halves of
        // I1, writes the high halves: disjoint words)
        e = launch_mul_strided(i + 1, I1, 2 * h, h, T3 + h, 2 * h, h, I1 + h, 2 * h, h, N, fam, s);
    }
    if (e != hipSuccess) return e;
    if (keep && keep->trees_only) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_copy, dim3(eblocks(n)), dim3(256), 0, s, F_in, n, p, Fr);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const u32 *F = Fr;
    u32 *kept = keep ? keep->rem : nullptr;
    // ---- remainder tree: R_k = F mod p, then levels k-1 .. L0; the
    // reversed top halves of a level's dividends come from the previous
    // level's k_div_finish (from F at the first level)
    // levels below `stop` (from level pin = min(k, kEvalFuse) down) run in
    // k_rem_descend, one workgroup per node of level pin
    const int pin = k < kEvalFuse ? k : kEvalFuse;
    const bool fuse = pin > L0 && !kept;
    const int stop = fuse ? pin : L0;
    const u32 *A = F;
    u32 *Rn = R0;
    if (k - 1 >= stop) {
        const int h = 1 << (k - 1);
        hipLaunchKernelGGL(k_div_rev_top, dim3(eblocks(n)), dim3(256), 0, s, F, h,
                           n >> (k - 1), T1);
        e = hipGetLastError();
    }
    for (int i = k - 1; i >= stop && e == hipSuccess; --i) {
        const int h = 1 << i;
        const int N = (int)(n >> i);                         // nodes of level i
        // qrev(j) = revtop(j) I_i(j) mod x^h
        e = launch_mul_strided(i + 1, T1, h, h, inv_at(i), h, h, T2, h, h, N, fam, s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_reverse, dim3(eblocks((int64_t)N * h)), dim3(256), 0, s, T2, h,
                           (int64_t)N, T1);
        // qm(j) = q(j) M_i(j) mod x^h (the leading x^h of M vanishes mod x^h)
        e = launch_mul_strided(i + 1, T1, h, h, node(i), h + 1, h, T2, h, h, N, fam, s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_div_finish, dim3(eblocks((int64_t)N * h)), dim3(256), 0, s, A, T2, h,
                           (int64_t)N, p, Rn, i > stop ? T1 : nullptr);
        e = hipGetLastError();
        if (e == hipSuccess && kept)
            e = hipMemcpyAsync(kept + (int64_t)i * n, Rn, sizeof(u32) * (size_t)n,
                               hipMemcpyDeviceToDevice, s);
        A = Rn;
        Rn = Rn == R0 ? R1 : R0;
    }
    if (e != hipSuccess) return e;
    if (fuse) {
        LevelOffsets lo{};
        for (int l = 0; l <= k && l < 32; ++l) lo.off[l] = tree_level_offset(k, l);
        hipLaunchKernelGGL(k_rem_descend, dim3((unsigned)(n >> pin)), dim3(kEvalParts << pin), 0, s, A,
                           pin, L0, tree, lo, inv, n, pts, plain_mod(p), values);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_horner, dim3(eblocks(n)), dim3(256), 0, s, A, 1 << L0, pts, n,
                       plain_mod(p), values);
    e = hipGetLastError();
    // kept remainders below L0 (level L0 is in kept, or F itself when k = L0)
    for (int i = L0 - 1; kept && i >= 0 && e == hipSuccess; --i) {
        const u32 *par = i + 1 == k ? F : kept + (int64_t)(i + 1) * n;
        const int64_t N = n >> i;
        const u32 *Mi = node(i);
        u32 *out = kept + (int64_t)i * n;
        switch (i) {
        case 3: hipLaunchKernelGGL(k_rem_small<8>, dim3(eblocks(N)), dim3(256), 0, s, par, Mi, N, p, out); break;
        case 2: hipLaunchKernelGGL(k_rem_small<4>, dim3(eblocks(N)), dim3(256), 0, s, par, Mi, N, p, out); break;
        case 1: hipLaunchKernelGGL(k_rem_small<2>, dim3(eblocks(N)), dim3(256), 0, s, par, Mi, N, p, out); break;
        default: hipLaunchKernelGGL(k_rem_small<1>, dim3(eblocks(N)), dim3(256), 0, s, par, Mi, N, p, out); break;
        }
        e = hipGetLastError();
    }
    return e;
}

}  // namespace cm
