// plain.hip -- schoolbook products modulo any odd modulus below 2^31.
//
// Reference: mulCUDA (src/opt_plain_mul.cu:258-384, include/opt_plain_mul.h:66),
// the plain multiply of the HPCS 2012 paper: a multiplication phase of
// BLOCK x BLOCK tiles into a (rows x (n + BLOCK - 1)) scratch matrix, then
// log2(rows) addition kernels and a final copy.  It needs no root of unity,
// so it serves moduli the NTT cannot, and it is the cheaper product for short
// operands (the reference multiplies by FFT only from 256 points up,
// src/stockham.cu:545).
//
// Here one wavefront computes 256 consecutive output coefficients in one
// pass with no scratch matrix: lane t owns outputs c0 + 4t .. c0 + 4t + 3 and
// accumulates their products in 64-bit registers.  B (the shorter operand)
// is swept four coefficients at a time; each step reads the four B values as
// one broadcast 128-bit LDS load and the four A values that enter the lane's
// window as one aligned 128-bit LDS load (lanes 16 bytes apart: no bank
// conflicts), then issues 16 v_mad_u64_u32.  Accumulators stay below 2^64 by
// folding acc -> hi * c + lo, c = 2^32 mod p (congruent mod p), after every
// four products.  A folded value is at most B1 = (2^32 - 1)(c + 1), and
// B1 + 4 (p-1)^2 < 2^64 for every p < 2^31: with k = floor(2^32 / p) >= 2 and
// c = 2^32 - k p the sum is 2^64 - k 2^32 p + 4 p^2 + O(p), below 2^64 when
// 4 p < k 2^32 (k = 2: p < 2^31; k >= 3: p <= 2^32 / k < k 2^30).  Each output
// is reduced to [0, p) once at the end (two folds, one Montgomery reduction,
// one Montgomery product by 2^64 mod p).
#include <cstring>
#include <vector>

#include "internal.h"

namespace cm {

constexpr int kPlainLanes = 64;
constexpr int kPlainR = 4;                            // outputs per lane
constexpr int kPlainChunk = kPlainLanes * kPlainR;    // outputs per workgroup
constexpr int kPlainJT = 512;                         // B coefficients per pass
constexpr int kPlainALds = kPlainChunk + kPlainJT + 8;

struct PlainChunk {
    int32_t item;   // descriptor index
    int32_t c0;     // first output coefficient
};

struct PlainMod {
    u32 p, np, c32, r2;   // modulus, -p^-1 mod 2^32, 2^32 mod p, 2^64 mod p
};

__device__ __forceinline__ u64 plain_fold(u64 acc, u32 c32) {
    return (u64)(u32)(acc >> 32) * c32 + (u32)acc;
}

// acc (any 64-bit value) -> acc mod p
__device__ __forceinline__ u32 plain_reduce(u64 acc, const PlainMod &m) {
    acc = plain_fold(plain_fold(acc, m.c32), m.c32);    // < 2^62 + 2^33
    const u32 q = (u32)acc * m.np;
    const u32 r = (u32)((acc + (u64)q * m.p) >> 32);    // acc / 2^32 mod p, < 2^32
    return cm_mul_tw(r, m.r2, m.p, m.np);               // r 2^64 / 2^32 = acc mod p
}

// Descriptor d: operands F + d.f (d.lf words) and G + d.g (d.lg words),
// output H + d.h (d.lh = lf + lg - 1 words).  The sweep runs over the shorter
// operand: d.logn != 0 marks a descriptor whose G operand is the longer one.
// Operands are canonical residues.  One workgroup (one wavefront) per chunk
// of 256 outputs.
__global__ __launch_bounds__(kPlainLanes) void k_plain_mul(const u32 *__restrict__ F,
                                                           const u32 *__restrict__ G,
                                                           u32 *__restrict__ H,
                                                           const MulDesc *__restrict__ descs,
                                                           const PlainChunk *__restrict__ chunks,
                                                           PlainMod m) {
    __shared__ __align__(16) u32 sa[kPlainALds];
    __shared__ __align__(16) u32 sb[kPlainJT];
    const PlainChunk ch = chunks[blockIdx.x];
    const MulDesc d = descs[ch.item];
    const bool sw = d.logn != 0;
    const u32 *A = sw ? G + d.g : F + d.f;
    const u32 *B = sw ? F + d.f : G + d.g;
    const int la = sw ? d.lg : d.lf, lb = sw ? d.lf : d.lg;
    const int t = threadIdx.x;
    const int c0 = ch.c0;
    // terms j of output c need 0 <= c - j < la and 0 <= j < lb
    int j_lo = c0 - (la - 1);
    j_lo = j_lo < 0 ? 0 : j_lo & ~3;
    int j_hi = c0 + kPlainChunk - 1;                      // last j that can contribute
    j_hi = j_hi < lb - 1 ? j_hi : lb - 1;
    u64 acc[kPlainR] = {0, 0, 0, 0};
    for (int js = j_lo; js <= j_hi; js += kPlainJT) {
        const int nj = (j_hi - js + 1 + 3) & ~3;          // multiple of 4
        const int jn = nj < kPlainJT ? nj : kPlainJT;     // terms this pass
        const int je = js + jn;
        // sa[x] = A[a_lo + x] for a_lo = c0 - je (zero outside [0, la)),
        // x < kPlainChunk + jn;  sb[k] = B[js + k] (zero beyond lb)
        const int a_lo = c0 - je;
        const int na = kPlainChunk + jn;
        __syncthreads();
        for (int x = t; x < na; x += kPlainLanes) {
            const int i = a_lo + x;
            sa[x] = (i >= 0 && i < la) ? A[i] : 0u;
        }
        for (int k = t; k < jn; k += kPlainLanes) {
            const int j = js + k;
            sb[k] = j < lb ? B[j] : 0u;
        }
        __syncthreads();
        // window W = A[c0 + 4t - j .. c0 + 4t - j + 3] at j = js
        uint4 w = *(const uint4 *)(sa + 4 * t + jn);
        for (int k = 0; k < jn; k += 4) {
            // entering values A[c0 + 4t - j - 4 .. c0 + 4t - j - 1], j = js + k
            const uint4 nw = *(const uint4 *)(sa + 4 * t + jn - k - 4);
            const uint4 bv = *(const uint4 *)(sb + k);
            const u32 x[8] = {nw.x, nw.y, nw.z, nw.w, w.x, w.y, w.z, w.w};
            const u32 b[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int r = 0; r < kPlainR; ++r) acc[r] += (u64)x[4 + r - s] * b[s];
            }
#pragma unroll
            for (int r = 0; r < kPlainR; ++r) acc[r] = plain_fold(acc[r], m.c32);
            w = nw;
        }
    }
    u32 *Hc = H + d.h;
    const int c = c0 + kPlainR * t;
    if (c + kPlainR <= d.lh) {
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
        for (int r = 0; r < kPlainR; ++r)
            if (c + r < d.lh) Hc[c + r] = plain_reduce(acc[r], m);
    }
}

static PlainMod plain_mod(u32 p) {
    PlainMod m;
    m.p = p;
    m.np = cm_host_npinv(p);
    m.c32 = (u32)(((u64)1 << 32) % p);
    m.r2 = (u32)((unsigned __int128)((u64)m.c32 * m.c32) % p);
    return m;
}

hipError_t launch_plain_list(const MulDesc *descs, int64_t count, const u32 *F, const u32 *G,
                             u32 *H, u32 p, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (count > INT32_MAX) return hipErrorInvalidValue;
    std::vector<PlainChunk> chunks;
    for (int64_t i = 0; i < count; ++i)
        for (int64_t c0 = 0; c0 < descs[i].lh; c0 += kPlainChunk)
            chunks.push_back(PlainChunk{(int32_t)i, (int32_t)c0});
    if (chunks.size() > (size_t)INT32_MAX) return hipErrorInvalidValue;
    const size_t desc_bytes = sizeof(MulDesc) * (size_t)count;
    const size_t chunk_bytes = sizeof(PlainChunk) * chunks.size();
    std::vector<char> blob(desc_bytes + chunk_bytes);
    memcpy(blob.data(), descs, desc_bytes);
    memcpy(blob.data() + desc_bytes, chunks.data(), chunk_bytes);
    char *dblob = (char *)workspace(17, blob.size(), s);
    if (!dblob) return hipErrorOutOfMemory;
    hipError_t e = stage_upload(blob.data(), blob.size(), dblob, s);
    if (e != hipSuccess) return e;
    const MulDesc *d_descs = (const MulDesc *)dblob;
    const PlainChunk *d_chunks = (const PlainChunk *)(dblob + desc_bytes);
    const PlainMod m = plain_mod(p);
    ProfScope ps(kProfPlainMul, s);
    const dim3 g((unsigned)chunks.size());
    hipLaunchKernelGGL(k_plain_mul, g, dim3(kPlainLanes), 0, s, F, G, H, d_descs, d_chunks, m);
    return hipGetLastError();
}

}  // namespace cm
