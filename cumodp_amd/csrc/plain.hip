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
// conflicts), then issues 16 v_mad_u64_u32 into 64-bit accumulators that
// are folded after every four products (accum.h).
#include <cstring>
#include <vector>

#include "accum.h"
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

// Descriptor d: operands F + d.f (d.lf words) and G + d.g (d.lg words),
// output H + d.h (d.lh = lf + lg - 1 words).  The sweep runs over the shorter
// operand: d.logn != 0 marks a descriptor whose G operand is the longer one.
// Operands are canonical residues.  One workgroup (one wavefront) per chunk
// of 256 outputs.  UNIFORM: every item has the shape of `one`, whose f, g, h
// fields are the strides between items (item = block / cpi, chunk = block %
// cpi); otherwise chunks[] and descs[] list them.
template <bool UNIFORM>
__global__ __launch_bounds__(kPlainLanes) void k_plain_mul(const u32 *__restrict__ F,
                                                           const u32 *__restrict__ G,
                                                           u32 *__restrict__ H,
                                                           const MulDesc *__restrict__ descs,
                                                           const PlainChunk *__restrict__ chunks,
                                                           MulDesc one, int cpi, PlainMod m) {
    __shared__ __align__(16) u32 sa[kPlainALds];
    __shared__ __align__(16) u32 sb[kPlainJT];
    PlainChunk ch;
    MulDesc d;
    if constexpr (UNIFORM) {
        const int64_t item = blockIdx.x / cpi;
        ch.item = 0;
        ch.c0 = (int32_t)(blockIdx.x % cpi) * kPlainChunk;
        d = one;
        d.f = item * one.f;
        d.g = item * one.g;
        d.h = item * one.h;
    } else {
        ch = chunks[blockIdx.x];
        d = descs[ch.item];
    }
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


hipError_t launch_plain_list(const MulDesc *descs, int64_t count, const u32 *F, const u32 *G,
                             u32 *H, u32 p, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (count > INT32_MAX) return hipErrorInvalidValue;
    int64_t nchunks = 0;
    for (int64_t i = 0; i < count; ++i) nchunks += (descs[i].lh + kPlainChunk - 1) / kPlainChunk;
    if (nchunks > INT32_MAX) return hipErrorInvalidValue;
    // descriptors then chunks, written straight into the pinned staging buffer
    const size_t desc_bytes = sizeof(MulDesc) * (size_t)count;
    const size_t bytes = desc_bytes + sizeof(PlainChunk) * (size_t)nchunks;
    char *dblob = (char *)workspace(17, bytes, s);
    if (!dblob) return hipErrorOutOfMemory;
    hipError_t e = stage_upload_fill(bytes, dblob, s, [&](char *blob) {
        memcpy(blob, descs, desc_bytes);
        PlainChunk *ch = (PlainChunk *)(blob + desc_bytes);
        for (int64_t i = 0; i < count; ++i)
            for (int32_t c0 = 0; c0 < descs[i].lh; c0 += kPlainChunk)
                *ch++ = PlainChunk{(int32_t)i, c0};
    });
    if (e != hipSuccess) return e;
    const size_t chunk_count = (size_t)nchunks;
    const MulDesc *d_descs = (const MulDesc *)dblob;
    const PlainChunk *d_chunks = (const PlainChunk *)(dblob + desc_bytes);
    const PlainMod m = plain_mod(p);
    ProfScope ps(kProfPlainMul, s);
    const dim3 g((unsigned)chunk_count);
    hipLaunchKernelGGL(k_plain_mul<false>, g, dim3(kPlainLanes), 0, s, F, G, H, d_descs,
                       d_chunks, MulDesc{}, 0, m);
    return hipGetLastError();
}

hipError_t launch_plain_strided(const u32 *F, int64_t fs, int lf, const u32 *G, int64_t gs,
                                int lg, u32 *H, int64_t hs, int lh, int count, u32 p,
                                hipStream_t s) {
    if (count <= 0 || lh <= 0) return hipSuccess;
    if (lf < 1 || lg < 1 || (int64_t)lh > (int64_t)lf + lg - 1) return hipErrorInvalidValue;
    const int cpi = (lh + kPlainChunk - 1) / kPlainChunk;
    if ((int64_t)count * cpi > INT32_MAX) return hipErrorInvalidValue;
    const MulDesc one{fs, gs, hs, lf, lg, lh, lg > lf ? 1 : 0};
    const PlainMod m = plain_mod(p);
    ProfScope ps(kProfPlainMul, s);
    hipLaunchKernelGGL(k_plain_mul<true>, dim3((unsigned)(count * cpi)), dim3(kPlainLanes), 0, s,
                       F, G, H, nullptr, nullptr, one, cpi, m);
    return hipGetLastError();
}

}  // namespace cm
