// Exclusive prefix sum with an optional fused gather, for gfx950
// (the contract of crates/sort/src/scan.rs:182-255).
//
// Three launches, each of which keeps all 256 CUs busy rather than a single
// workgroup: block reduce (4096 cells per block, gather fused) -> scan of the
// block sums (one workgroup, 4096 sums per sweep) -> per-block scan + apply.
// In place is allowed (gids == nullptr and counts == offsets).
#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

constexpr uint32_t NT = 256;
constexpr uint32_t EPT = 16;
constexpr uint32_t BLK = NT * EPT;
static_assert(BLK == kScanBlock, "scan block");

// Gathered cell i: counts[gids[i]]; with segments of `seg` cells, gids hold
// in-segment indices into the segment's own counts.
__device__ __forceinline__ size_t gather_index(const uint32_t* gids, size_t i, uint32_t seg) {
    if (!gids) return i;
    return seg ? (i / seg) * seg + gids[i] : gids[i];
}

__global__ __launch_bounds__(NT) void k_scan_reduce(const uint32_t* __restrict__ gids,
                                                    const uint32_t* __restrict__ counts,
                                                    uint32_t n, uint32_t* __restrict__ sums,
                                                    uint32_t seg) {
    __shared__ uint32_t tmp[NT / 64 + 1];
    const size_t base = (size_t)blockIdx.x * BLK + threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t i = base + (size_t)e * NT;
        if (i < n) s += counts[gather_index(gids, i, seg)];
    }
    uint32_t total;
    block_excl_scan_u32<NT>(s, tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(NT) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb,
                                                  uint32_t* __restrict__ total_out) {
    __shared__ uint32_t tmp[NT / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += BLK) {
        uint32_t v[EPT];
        uint32_t s = 0;
        uint32_t first = c0 + threadIdx.x * EPT;
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            v[e] = first + e < nb ? sums[first + e] : 0u;
            s += v[e];
        }
        uint32_t total;
        uint32_t run = carry + block_excl_scan_u32<NT>(s, tmp, &total);
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            if (first + e < nb) sums[first + e] = run;
            run += v[e];
        }
        carry += total;
    }
    if (total_out && threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(NT) void k_scan_apply(const uint32_t* __restrict__ gids,
                                                   const uint32_t* counts, uint32_t n,
                                                   const uint32_t* __restrict__ sums,
                                                   uint32_t* offsets, uint32_t seg) {
    __shared__ uint32_t tmp[NT / 64 + 1];
    const size_t first = (size_t)blockIdx.x * BLK + (size_t)threadIdx.x * EPT;
    uint32_t v[EPT];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t i = first + e;
        v[e] = i < n ? counts[gather_index(gids, i, seg)] : 0u;
        s += v[e];
    }
    uint32_t total;
    uint32_t run = sums[blockIdx.x] + block_excl_scan_u32<NT>(s, tmp, &total);
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t i = first + e;
        if (i < n) offsets[i] = run;
        run += v[e];
    }
}

}  // namespace

hipError_t launch_scan(hipStream_t s, const uint32_t* gids, const uint32_t* counts,
                       uint32_t* offsets, uint32_t n, uint32_t* block_sums, uint32_t seg,
                       uint32_t* total_out) {
    if (n == 0) return hipSuccess;
    uint32_t nb = div_up(n, BLK);
    k_scan_reduce<<<nb, NT, 0, s>>>(gids, counts, n, block_sums, seg);
    k_scan_sums<<<1, NT, 0, s>>>(block_sums, nb, total_out);
    k_scan_apply<<<nb, NT, 0, s>>>(gids, counts, n, block_sums, offsets, seg);
    return hipGetLastError();
}

}  // namespace sfk
