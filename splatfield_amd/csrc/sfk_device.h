// Wavefront (64-lane) and workgroup helpers for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfk {

// Workgroup barrier ordering LDS only: unlike __syncthreads it does not wait
// for this wave's outstanding global loads (a prefetch stays in flight).  For
// code where no thread reads global memory another thread of the workgroup
// wrote before the barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum across the 64 lanes with DPP moves, valid in lane 63 only: butterflies
// inside each 16-lane row (quad_perm xor 1, xor 2, row_half_mirror,
// row_mirror leave every lane of a row with the row sum), then row_bcast15
// into rows 1 and 3 and row_bcast31 into rows 2 and 3.  Six VALU adds with a
// DPP source operand instead of six LDS-crossbar shuffles.
#define SFK_DPP_I(x, ctrl, rmask) __builtin_amdgcn_update_dpp(0, (x), (ctrl), (rmask), 0xf, false)

__device__ __forceinline__ uint32_t wave_sum_last_u32(uint32_t v) {
    int x = (int)v;
    x += SFK_DPP_I(x, 0xb1, 0xf);   // quad_perm [1,0,3,2]
    x += SFK_DPP_I(x, 0x4e, 0xf);   // quad_perm [2,3,0,1]
    x += SFK_DPP_I(x, 0x141, 0xf);  // row_half_mirror
    x += SFK_DPP_I(x, 0x140, 0xf);  // row_mirror
    x += SFK_DPP_I(x, 0x142, 0xa);  // row_bcast15 -> rows 1, 3
    x += SFK_DPP_I(x, 0x143, 0xc);  // row_bcast31 -> rows 2, 3
    return (uint32_t)x;
}

__device__ __forceinline__ float wave_sum_last_f32(float v) {
#define SFK_DPP_F(x, ctrl, rmask) \
    __builtin_bit_cast(float, SFK_DPP_I(__builtin_bit_cast(int, (x)), (ctrl), (rmask)))
    v += SFK_DPP_F(v, 0xb1, 0xf);
    v += SFK_DPP_F(v, 0x4e, 0xf);
    v += SFK_DPP_F(v, 0x141, 0xf);
    v += SFK_DPP_F(v, 0x140, 0xf);
    v += SFK_DPP_F(v, 0x142, 0xa);
    v += SFK_DPP_F(v, 0x143, 0xc);
#undef SFK_DPP_F
    return v;
}

// Eight sums over the 64 lanes at once (a transposed butterfly).  Lane l
// holds v[0..7]; afterwards lane l returns the wave sum of v[l >> 3], so each
// sum sits in the 8 lanes 8s..8s+7.  The two cross-row halvings use the
// gfx950 v_permlane32_swap / v_permlane16_swap moves, the cross-half one a
// row_ror:8 DPP move, the last three the in-row butterflies; 22 VALU
// operations for eight sums instead of 8 x 12.
template <typename T>
__device__ __forceinline__ T wave_sum8_transposed(const T (&v)[8]) {
    static_assert(sizeof(T) == 4, "32-bit lanes");
    auto bits = [](T x) { return __builtin_bit_cast(uint32_t, x); };
    auto val = [](uint32_t x) { return __builtin_bit_cast(T, x); };
    // Upper 32 lanes of the first operand <-> lower 32 lanes of the second:
    // u[i] = (lanes 0-31: v[i], lanes 32-63: v[i + 4]) summed over l, l ^ 32.
    T u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        auto r = __builtin_amdgcn_permlane32_swap(bits(v[i]), bits(v[i + 4]), false, false);
        u[i] = val(r[0]) + val(r[1]);
    }
    // Odd rows of the first operand <-> even rows of the second: row r of
    // w[i] holds v[2r + i] summed over l, l ^ 16, l ^ 32, l ^ 48.
    T w[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        auto r = __builtin_amdgcn_permlane16_swap(bits(u[i]), bits(u[i + 2]), false, false);
        w[i] = val(r[0]) + val(r[1]);
    }
    // Lanes with bit 3 clear keep w[0] (value 2r) and take the partner's
    // w[0] across row_ror:8; lanes with bit 3 set keep w[1] (value 2r + 1).
    const bool hi = (lane_id() & 8u) != 0;
    T keep = hi ? w[1] : w[0];
    T send = hi ? w[0] : w[1];
    T x = keep + val((uint32_t)SFK_DPP_I((int)bits(send), 0x128, 0xf));  // row_ror:8
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0xb1, 0xf));   // quad_perm [1,0,3,2]
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0x4e, 0xf));   // quad_perm [2,3,0,1]
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0x141, 0xf));  // row_half_mirror
    return x;
}

// Four sums over the 64 lanes at once: the same transposed butterfly as
// wave_sum8_transposed stopped one level earlier.  Lane l returns the wave
// sum of v[l >> 4] (each sum in the 16 lanes of row l >> 4); 10 VALU
// operations for four sums.
template <typename T>
__device__ __forceinline__ T wave_sum4_transposed(const T (&v)[4]) {
    static_assert(sizeof(T) == 4, "32-bit lanes");
    auto bits = [](T x) { return __builtin_bit_cast(uint32_t, x); };
    auto val = [](uint32_t x) { return __builtin_bit_cast(T, x); };
    // u[i]: lanes 0-31 v[i], lanes 32-63 v[i + 2], summed over l, l ^ 32.
    T u[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        auto r = __builtin_amdgcn_permlane32_swap(bits(v[i]), bits(v[i + 2]), false, false);
        u[i] = val(r[0]) + val(r[1]);
    }
    // Row r holds v[r] summed over l, l ^ 16, l ^ 32, l ^ 48.
    auto r = __builtin_amdgcn_permlane16_swap(bits(u[0]), bits(u[1]), false, false);
    T x = val(r[0]) + val(r[1]);
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0x128, 0xf));  // row_ror:8
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0x141, 0xf));  // row_half_mirror
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0x4e, 0xf));   // quad_perm [2,3,0,1]
    x += val((uint32_t)SFK_DPP_I((int)bits(x), 0xb1, 0xf));   // quad_perm [1,0,3,2]
    return x;
}

// 64 x 64 bit-matrix transpose across a wavefront: lane l holds row l (bit j
// = element (l, j)) and receives column l (bit i = element (i, l)).  Stage k
// (k = 32, 16, 8, 4, 2, 1) swaps bit k of the lane index with bit k of the
// bit index: element (l, j) with (l & k) != (j & k) trades places with
// (l ^ k, j ^ k).  Lanes move by v_permlane32_swap (k = 32),
// v_permlane16_swap (16), DPP row_ror:8 (8), a bpermute (4) and DPP
// quad_perm (2, 1); bits by v_perm byte selects (k >= 8) or a rotate and a
// bit-field insert (k < 8).  About 30 VALU operations for the 4096 bits.
struct TransposeLanes {
    uint32_t sel16, sel8;     // v_perm selectors of the byte stages
    uint32_t rot[3], msk[3];  // rotate amount and insert mask of stages 4, 2, 1
};

__device__ __forceinline__ TransposeLanes transpose_lanes() {
    const uint32_t l = lane_id();
    TransposeLanes c;
    c.sel16 = (l & 16u) ? 0x07060302u : 0x05040100u;
    c.sel8 = (l & 8u) ? 0x03070105u : 0x06020400u;
    // Lanes with the stage bit clear take the partner's low field into their
    // high field (rotate left by k); the others the partner's high field into
    // their low field (rotate right by k).
    c.rot[0] = (l & 4u) ? 4u : 28u;
    c.msk[0] = (l & 4u) ? 0x0f0f0f0fu : 0xf0f0f0f0u;
    c.rot[1] = (l & 2u) ? 2u : 30u;
    c.msk[1] = (l & 2u) ? 0x33333333u : 0xccccccccu;
    c.rot[2] = (l & 1u) ? 1u : 31u;
    c.msk[2] = (l & 1u) ? 0x55555555u : 0xaaaaaaaau;
    return c;
}

__device__ __forceinline__ uint32_t transpose_insert(uint32_t w, uint32_t partner, uint32_t rot,
                                                     uint32_t msk) {
    const uint32_t t = __builtin_amdgcn_alignbit(partner, partner, rot);  // rotate right
    return (t & msk) | (w & ~msk);
}

__device__ __forceinline__ uint64_t wave_transpose64(uint64_t row, const TransposeLanes& c) {
    uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
    // k = 32: lanes 0-31 take their partner's low word as the high word,
    // lanes 32-63 their partner's high word as the low word.
    {
        auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    }
    // k = 16: permlane16_swap(w, w) leaves (w, partner) in (r[0], r[1]) on
    // even rows and (partner, w) on odd rows.
    {
        auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        lo = __builtin_amdgcn_perm(a[1], a[0], c.sel16);
        auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        hi = __builtin_amdgcn_perm(b[1], b[0], c.sel16);
    }
    // k = 8: row_ror:8 is lane ^ 8.
    {
        const uint32_t plo = (uint32_t)SFK_DPP_I((int)lo, 0x128, 0xf);
        const uint32_t phi = (uint32_t)SFK_DPP_I((int)hi, 0x128, 0xf);
        lo = __builtin_amdgcn_perm(plo, lo, c.sel8);
        hi = __builtin_amdgcn_perm(phi, hi, c.sel8);
    }
    // k = 4
    {
        const uint32_t plo = (uint32_t)__shfl_xor((int)lo, 4, 64);
        const uint32_t phi = (uint32_t)__shfl_xor((int)hi, 4, 64);
        lo = transpose_insert(lo, plo, c.rot[0], c.msk[0]);
        hi = transpose_insert(hi, phi, c.rot[0], c.msk[0]);
    }
    // k = 2: quad_perm [2,3,0,1]; k = 1: quad_perm [1,0,3,2].
    {
        const uint32_t plo = (uint32_t)SFK_DPP_I((int)lo, 0x4e, 0xf);
        const uint32_t phi = (uint32_t)SFK_DPP_I((int)hi, 0x4e, 0xf);
        lo = transpose_insert(lo, plo, c.rot[1], c.msk[1]);
        hi = transpose_insert(hi, phi, c.rot[1], c.msk[1]);
    }
    {
        const uint32_t plo = (uint32_t)SFK_DPP_I((int)lo, 0xb1, 0xf);
        const uint32_t phi = (uint32_t)SFK_DPP_I((int)hi, 0xb1, 0xf);
        lo = transpose_insert(lo, plo, c.rot[2], c.msk[2]);
        hi = transpose_insert(hi, phi, c.rot[2], c.msk[2]);
    }
    return ((uint64_t)hi << 32) | lo;
}

// Lanes of the wave (of those with `live` set) whose digit d agrees with this
// lane's in bits [0, W) (match-any): one ballot per bit.  With the lane's bit
// spread to all-ones or all-zeros m, each half of the mask updates as
// peers &= ~(ballot ^ m), which gfx950 does in one v_bitop3.  Digit bits at
// or above a digit's real width are zero in every lane and match trivially.
template <uint32_t W>
__device__ __forceinline__ uint64_t match_peers(uint32_t d, bool live) {
    const uint64_t l = __ballot(live);
    uint32_t lo = (uint32_t)l, hi = (uint32_t)(l >> 32);
#pragma unroll
    for (uint32_t b = 0; b < W; ++b) {
        const uint32_t m = (uint32_t)((int32_t)(d << (31u - b)) >> 31);
        const uint64_t bal = __ballot(m != 0u);
        lo &= ~((uint32_t)bal ^ m);
        hi &= ~((uint32_t)(bal >> 32) ^ m);
    }
    return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum across the 64 lanes of a wave with DPP moves: a
// Hillis-Steele scan inside each 16-lane row (row_shr 1, 2, 4, 8; lanes
// shifted in from outside the row read 0), then row_bcast15 adds row 0's
// total into row 1 and row 2's into row 3, and row_bcast31 adds rows 0-1's
// total into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    int x = (int)v;
    x += SFK_DPP_I(x, 0x111, 0xf);  // row_shr:1
    x += SFK_DPP_I(x, 0x112, 0xf);  // row_shr:2
    x += SFK_DPP_I(x, 0x114, 0xf);  // row_shr:4
    x += SFK_DPP_I(x, 0x118, 0xf);  // row_shr:8
    x += SFK_DPP_I(x, 0x142, 0xa);  // row_bcast15 -> rows 1, 3
    x += SFK_DPP_I(x, 0x143, 0xc);  // row_bcast31 -> rows 2, 3
    return (uint32_t)x;
}

// Exclusive prefix sum over a workgroup of NT threads (NT a multiple of 64,
// at most 1024); `tmp` holds NT/64 + 1 words of LDS.  Returns the exclusive
// prefix; *total receives the workgroup sum.  Contains barriers: every thread
// of the group must call it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* tmp,
                                                        uint32_t* total) {
    constexpr int NW = NT / 64;
    uint32_t lane = lane_id();
    uint32_t w = threadIdx.x / 64;
    uint32_t inc = wave_incl_scan_u32(v);
    if (lane == 63) tmp[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < NW; ++i) {
            uint32_t c = tmp[i];
            tmp[i] = run;
            run += c;
        }
        tmp[NW] = run;
    }
    __syncthreads();
    uint32_t r = tmp[w] + inc - v;
    *total = tmp[NW];
    __syncthreads();
    return r;
}

// Spread consecutive workgroup ids of a 1D grid over the 8 XCDs of the chip
// so that each XCD (and its L2) works on a contiguous range of logical ids:
// the dispatcher assigns workgroup i to XCD i % 8.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    uint32_t xcd = bid & 7u, q = bid >> 3;
    uint32_t per = nblocks >> 3, rem = nblocks & 7u;
    uint32_t base = xcd * per + (xcd < rem ? xcd : rem);
    return base + q;
}

}  // namespace sfk
