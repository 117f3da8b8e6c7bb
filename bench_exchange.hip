// bench_exchange.hip -- measurement probe run by bench_exchange.py (not part
// of the library): what the element exchanges between the register-resident
// rounds of an NTT cost when they go through LDS, and what they cost as
// cross-lane operations (DPP, v_permlane16/32_swap) inside one wavefront.
//
// Each wavefront holds 1024-point vectors mod p, 16 words per lane, and runs
// the forward transform (decimation in frequency, the library's canonical
// Montgomery butterflies) as three rounds of in-register stages, like the
// library's radix-16 rounds:
//   round A: element bits 9..6 (distances 512 .. 64),
//   round B: element bits 5..2 (32 .. 4),
//   round C: element bits 1..0 (2, 1).
// A "map" says which element bit each register-index bit and lane bit holds.
// Between the rounds the data changes map:
//   k_probe_lds:   through the wave's LDS region (32-bit stores, then 32-bit
//                  loads before round B and 128-bit loads before round C);
//   k_probe_xt:    cross-lane transposes: lane bits 5 and 4 trade places
//                  with register bits by v_permlane32_swap / v_permlane16_swap
//                  (one instruction per register pair), lane bits 3 and 2 by
//                  one bank-masked DPP move per register, lane bits 1 and 0
//                  by a DPP move and a select per register;
//   k_probe_xb:    no transpose: round B and C stages run across lanes, each
//                  lane fetching its partner's word (permlane swap or DPP)
//                  and computing both outputs of the butterfly, keeping its
//                  own (the textbook "DPP butterfly").
// All three apply the same transform; `reps` transforms run back to back on
// the registers between one load and one store, so the loop times the
// arithmetic and the exchanges only.  Each next repetition reads the
// registers in its starting map, whatever map the last one ended in: the
// work is the same, and only reps == 1 is compared against a DFT.
#include <hip/hip_runtime.h>

#include "cumodp_amd/csrc/ntt_core.h"

namespace {

constexpr int kWaves = 4;     // wavefronts per workgroup
constexpr int kRegs = 16;     // words per lane
constexpr int kVec = 1024;    // points per wavefront

// element bit held by register-index bit i (reg) and lane bit i (lane)
template <int ID> struct MapT;
template <> struct MapT<0> {  // e = lane + 64 k
    static constexpr int reg[4] = {6, 7, 8, 9}, lane[6] = {0, 1, 2, 3, 4, 5};
};
template <> struct MapT<1> {  // round B
    static constexpr int reg[4] = {2, 3, 4, 5}, lane[6] = {0, 1, 6, 7, 8, 9};
};
template <> struct MapT<2> {  // round C after cross-lane transposes
    static constexpr int reg[4] = {0, 1, 4, 5}, lane[6] = {2, 3, 6, 7, 8, 9};
};
template <> struct MapT<3> {  // round C after LDS: four consecutive words per register quad
    static constexpr int reg[4] = {0, 1, 8, 9}, lane[6] = {2, 3, 4, 5, 6, 7};
};

template <int ID>
__device__ __forceinline__ int elem(int lane, int k) {
    int e = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) e |= ((k >> i) & 1) << MapT<ID>::reg[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) e |= ((lane >> i) & 1) << MapT<ID>::lane[i];
    return e;
}

// DIF stage on the element bit held by register-index bit RB: pairs
// (k, k + 2^RB); twiddle w_{2d}^(e mod d) of the lower element e, d = 2^bit,
// from the level table (level M at offset M/2 - 1)
template <int ID, int RB>
__device__ __forceinline__ void stage_reg(u32 (&v)[kRegs], int lane, const u32 *lv, Mod md) {
    constexpr int d = 1 << MapT<ID>::reg[RB];
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        if ((k >> RB) & 1) continue;
        const u32 w = lv[d - 1 + (elem<ID>(lane, k) & (d - 1))];
        cm_dif_bfly(v[k], v[k | (1 << RB)], w, md);
    }
}

// the four stages of a round on register bits 3..0, highest element bit first
template <int ID>
__device__ __forceinline__ void round4(u32 (&v)[kRegs], int lane, const u32 *lv, Mod md) {
    stage_reg<ID, 3>(v, lane, lv, md);
    stage_reg<ID, 2>(v, lane, lv, md);
    stage_reg<ID, 1>(v, lane, lv, md);
    stage_reg<ID, 0>(v, lane, lv, md);
}

// LDS placement: 4 pad words after every 64, so every map's 32-bit access
// and the 128-bit access of map 3 reach distinct banks
__device__ __forceinline__ int phys(int x) { return x + 4 * (x >> 6); }
constexpr int kLdsWords = kVec + 4 * (kVec >> 6);

// order the wave's LDS accesses around an exchange (one wavefront's LDS
// operations complete in issue order)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int A, int B>
__device__ __forceinline__ void lds_exchange(u32 (&v)[kRegs], u32 *S, int lane) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) S[phys(elem<A>(lane, k))] = v[k];
    wave_sync();
    if constexpr (B == 3) {
#pragma unroll
        for (int q = 0; q < kRegs / 4; ++q) {
            const uint4 t = *(const uint4 *)(S + phys(elem<B>(lane, 4 * q)));
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kRegs; ++k) v[k] = S[phys(elem<B>(lane, k))];
    }
    wave_sync();
}

// word of lane (lane ^ 2^LB) for LB <= 3 (DPP inside a row of 16 lanes)
template <int LB>
__device__ __forceinline__ u32 dpp_xor(u32 x) {
    if constexpr (LB == 0) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
    else if constexpr (LB == 1) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else if constexpr (LB == 3) return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, true); // row_ror:8
    else return 0;
}

// Swap register-index bit RB with lane bit LB: for each pair (a, b = a +
// 2^RB), lanes whose bit LB is 0 keep a[lane] and receive a[lane + 2^LB] in
// b; lanes whose bit LB is 1 keep b[lane] and receive b[lane - 2^LB] in a.
template <int RB, int LB>
__device__ __forceinline__ void xpose_bit(u32 (&v)[kRegs], int lane) {
    const bool hi = (lane >> LB) & 1;
#pragma unroll
    for (int a = 0; a < kRegs; ++a) {
        if ((a >> RB) & 1) continue;
        const int b = a | (1 << RB);
        if constexpr (LB == 5) {
            const auto r = __builtin_amdgcn_permlane32_swap(v[a], v[b], false, false);
            v[a] = r[0];
            v[b] = r[1];
        } else if constexpr (LB == 4) {
            const auto r = __builtin_amdgcn_permlane16_swap(v[a], v[b], false, false);
            v[a] = r[0];
            v[b] = r[1];
        } else if constexpr (LB >= 2) {
            // DPP bank masks select the lanes: a bank is four lanes of a
            // 16-lane row, so lane bit 2 is bank bit 0 and lane bit 3 bank
            // bit 1.  a keeps its own word where bit LB is 0 and takes
            // b[lane - 2^LB] (row_shr) where it is 1; b the other way round.
            constexpr int shl = 0x100 | (1 << LB), shr = 0x110 | (1 << LB);
            constexpr int hib = LB == 2 ? 0xA : 0xC, lob = LB == 2 ? 0x5 : 0x3;
            const u32 na = __builtin_amdgcn_update_dpp(v[a], v[b], shr, 0xF, hib, false);
            const u32 nb = __builtin_amdgcn_update_dpp(v[b], v[a], shl, 0xF, lob, false);
            v[a] = na;
            v[b] = nb;
        } else {
            const u32 pa = dpp_xor<LB>(v[a]), pb = dpp_xor<LB>(v[b]);
            v[a] = hi ? pb : v[a];
            v[b] = hi ? v[b] : pa;
        }
    }
}

// DIF stage across lanes on lane bit LB (distance d = 2^LB, twiddle index
// lane mod d): every lane forms both outputs of its pair and keeps its own
template <int LB>
__device__ __forceinline__ void stage_xlane(u32 (&v)[kRegs], int lane, const u32 *lv, Mod md) {
    constexpr int d = 1 << LB;
    const bool hi = (lane >> LB) & 1;
    const u32 w = lv[d - 1 + (lane & (d - 1))];
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        u32 x, y;
        if constexpr (LB == 5) {
            const auto r = __builtin_amdgcn_permlane32_swap(v[k], v[k], false, false);
            x = r[0];
            y = r[1];
        } else if constexpr (LB == 4) {
            const auto r = __builtin_amdgcn_permlane16_swap(v[k], v[k], false, false);
            x = r[0];
            y = r[1];
        } else {
            u32 o;
            if constexpr (LB == 2) o = __builtin_amdgcn_ds_swizzle(v[k], (4 << 10) | 0x1F);  // xor 4
            else o = dpp_xor<LB>(v[k]);
            x = hi ? o : v[k];
            y = hi ? v[k] : o;
        }
        const u32 s = cm_reduce2p(x + y, md.p);
        const u32 t = cm_mul_signed(x - y, w, md.p, md.np);
        v[k] = hi ? t : s;
    }
}

template <int ID>
__device__ __forceinline__ void load_vec(u32 (&v)[kRegs], const u32 *base, int lane) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) v[k] = base[elem<ID>(lane, k)];
}

template <int ID>
__device__ __forceinline__ void store_vec(const u32 (&v)[kRegs], u32 *base, int lane) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) base[elem<ID>(lane, k)] = v[k];
}

}  // namespace

extern "C" __global__ __launch_bounds__(64 * kWaves) void k_probe_lds(u32 *X, const u32 *lv,
                                                                     Mod md, int reps) {
    __shared__ __align__(16) u32 sh[kWaves][kLdsWords];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u32 *base = X + ((int64_t)blockIdx.x * kWaves + w) * kVec;
    u32 v[kRegs];
    load_vec<0>(v, base, lane);
    for (int r = 0; r < reps; ++r) {
        round4<0>(v, lane, lv, md);
        lds_exchange<0, 1>(v, sh[w], lane);
        round4<1>(v, lane, lv, md);
        lds_exchange<1, 3>(v, sh[w], lane);
        stage_reg<3, 1>(v, lane, lv, md);
        stage_reg<3, 0>(v, lane, lv, md);
    }
    store_vec<3>(v, base, lane);
}

extern "C" __global__ __launch_bounds__(64 * kWaves) void k_probe_xt(u32 *X, const u32 *lv,
                                                                    Mod md, int reps) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u32 *base = X + ((int64_t)blockIdx.x * kWaves + w) * kVec;
    u32 v[kRegs];
    load_vec<0>(v, base, lane);
    for (int r = 0; r < reps; ++r) {
        round4<0>(v, lane, lv, md);
        // map 0 -> 1: register bits 3..0 (e9..e6) trade with lane bits 5..2
        xpose_bit<3, 5>(v, lane);
        xpose_bit<2, 4>(v, lane);
        xpose_bit<1, 3>(v, lane);
        xpose_bit<0, 2>(v, lane);
        round4<1>(v, lane, lv, md);
        // map 1 -> 2: register bits 1, 0 (e3, e2) trade with lane bits 1, 0
        xpose_bit<1, 1>(v, lane);
        xpose_bit<0, 0>(v, lane);
        stage_reg<2, 1>(v, lane, lv, md);
        stage_reg<2, 0>(v, lane, lv, md);
    }
    store_vec<2>(v, base, lane);
}

extern "C" __global__ __launch_bounds__(64 * kWaves) void k_probe_xb(u32 *X, const u32 *lv,
                                                                    Mod md, int reps) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u32 *base = X + ((int64_t)blockIdx.x * kWaves + w) * kVec;
    u32 v[kRegs];
    load_vec<0>(v, base, lane);
    for (int r = 0; r < reps; ++r) {
        round4<0>(v, lane, lv, md);
        stage_xlane<5>(v, lane, lv, md);
        stage_xlane<4>(v, lane, lv, md);
        stage_xlane<3>(v, lane, lv, md);
        stage_xlane<2>(v, lane, lv, md);
        stage_xlane<1>(v, lane, lv, md);
        stage_xlane<0>(v, lane, lv, md);
    }
    store_vec<0>(v, base, lane);
}

// C ABI for bench_exchange.py: `blocks` workgroups of four wavefronts on
// X (blocks * 4096 words), `reps` transforms each, on the null stream.
extern "C" int probe_launch(int variant, unsigned *X, const unsigned *lv, unsigned p,
                            unsigned np, int reps, int blocks) {
    const Mod md{p, np};
    const dim3 grid(blocks), block(64 * kWaves);
    switch (variant) {
    case 0: hipLaunchKernelGGL(k_probe_lds, grid, block, 0, 0, X, lv, md, reps); break;
    case 1: hipLaunchKernelGGL(k_probe_xt, grid, block, 0, 0, X, lv, md, reps); break;
    case 2: hipLaunchKernelGGL(k_probe_xb, grid, block, 0, 0, X, lv, md, reps); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}
