// Stable LSD radix argsort for gfx950 (the contract of
// crates/sort/src/sort.rs:302-381, re-designed for 64-lane wavefronts).
//
// One pass over a `width`-bit digit (<= 8) is three launches:
//   upsweep   per-block digit histogram (LDS, per-wave sub-histograms),
//             written digit-major hist[d * max_blocks + b];
//   digscan   one workgroup per digit: exclusive scan of its row over the
//             blocks, row total into totals[d];
//   scatter   per block: stable rank of every key inside the block with
//             64-lane ballot match-any (one ballot per digit bit — a 64-bit
//             ballot word, no 32-lane assumption as in sort.rs:57-87), LDS
//             staging in sorted order, then coalesced runs to the global
//             offset totals-prefix + hist-prefix.
// Block = 256 threads x 16 keys; wave w owns keys [w*1024, (w+1)*1024) of
// its block, item i lane l -> w*1024 + i*64 + l, so ranks follow input order.
// The element count may come from device memory (min(*d_n, n)): blocks past
// the live count exit at once, so a capacity-sized grid needs no host sync.
#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

constexpr uint32_t NT = kSortThreads;
constexpr uint32_t EPT = kSortEpt;
constexpr uint32_t KPB = kSortBlock;
constexpr uint32_t NW = NT / 64;

__device__ __forceinline__ uint32_t live_count(uint32_t n, const uint32_t* d_n) {
    if (d_n) {
        uint32_t m = *d_n;
        return m < n ? m : n;
    }
    return n;
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

template <typename K, int RADIX>
__global__ __launch_bounds__(NT) void k_upsweep(const K* __restrict__ keys, uint32_t n_max,
                                                const uint32_t* __restrict__ d_n, uint32_t shift,
                                                uint32_t mask, uint32_t* __restrict__ hist,
                                                uint32_t max_blocks) {
    __shared__ uint32_t h[NW][RADIX];
    const uint32_t n = live_count(n_max, d_n);
    const uint32_t nb = (n + KPB - 1) / KPB;
    const uint32_t b = blockIdx.x;
    if (b >= nb) return;
    const uint32_t w = threadIdx.x / 64;
    for (uint32_t i = threadIdx.x; i < NW * RADIX; i += NT) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t base = (size_t)b * KPB + threadIdx.x;
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t idx = base + (size_t)e * NT;
        if (idx < n) atomicAdd(&h[w][digit_of(keys[idx], shift, mask)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < RADIX; d += NT) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t k = 0; k < NW; ++k) t += h[k][d];
        hist[(size_t)d * max_blocks + b] = t;
    }
}

// One workgroup per digit: exclusive scan of hist[d][0..nb) in place.
template <int RADIX>
__global__ __launch_bounds__(NT) void k_digscan(uint32_t* __restrict__ hist, uint32_t max_blocks,
                                                uint32_t n_max, const uint32_t* __restrict__ d_n,
                                                uint32_t* __restrict__ totals) {
    __shared__ uint32_t tmp[NW + 1];
    const uint32_t n = live_count(n_max, d_n);
    const uint32_t nb = (n + KPB - 1) / KPB;
    const uint32_t d = blockIdx.x;
    uint32_t* row = hist + (size_t)d * max_blocks;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += NT * EPT) {
        uint32_t v[EPT];
        uint32_t s = 0;
        uint32_t first = c0 + threadIdx.x * EPT;
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            uint32_t idx = first + e;
            v[e] = idx < nb ? row[idx] : 0u;
            s += v[e];
        }
        uint32_t total;
        uint32_t ex = block_excl_scan_u32<NT>(s, tmp, &total);
        uint32_t run = carry + ex;
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            uint32_t idx = first + e;
            if (idx < nb) row[idx] = run;
            run += v[e];
        }
        carry += total;
    }
    if (threadIdx.x == 0) totals[d] = carry;
}

template <typename K, int RADIX>
__global__ __launch_bounds__(NT) void k_scatter(const K* __restrict__ keys_in,
                                                const uint32_t* __restrict__ vals_in,
                                                K* __restrict__ keys_out,
                                                uint32_t* __restrict__ vals_out, uint32_t n_max,
                                                const uint32_t* __restrict__ d_n, uint32_t shift,
                                                uint32_t width, const uint32_t* __restrict__ hist,
                                                uint32_t max_blocks,
                                                const uint32_t* __restrict__ totals) {
    __shared__ K s_keys[KPB];
    __shared__ uint32_t s_vals[KPB];
    __shared__ uint32_t s_wh[NW][RADIX];  // per-wave digit counters, then bases
    __shared__ uint32_t s_gofs[RADIX];    // global position minus local position
    __shared__ uint32_t s_tmp[NW + 1];

    const uint32_t n = live_count(n_max, d_n);
    const uint32_t nb = (n + KPB - 1) / KPB;
    const uint32_t b = blockIdx.x;
    if (b >= nb) return;
    const uint32_t t = threadIdx.x, w = t / 64, lane = lane_id();
    const uint32_t mask = (1u << width) - 1u;
    const uint32_t nbins = 1u << width;
    const size_t blk0 = (size_t)b * KPB;
    const uint32_t live = (uint32_t)((n - blk0) < KPB ? (n - blk0) : KPB);

    for (uint32_t i = t; i < NW * RADIX; i += NT) (&s_wh[0][0])[i] = 0;

    // Global digit base: exclusive scan of the digit totals (RADIX <= NT).
    {
        uint32_t tot = t < nbins ? totals[t] : 0u;
        uint32_t all;
        uint32_t ex = block_excl_scan_u32<NT>(tot, s_tmp, &all);
        if (t < (uint32_t)RADIX) s_gofs[t] = t < nbins ? ex + hist[(size_t)t * max_blocks + b] : 0u;
    }
    // (the scan's barriers also order the s_wh clear)

    K key[EPT];
    uint32_t val[EPT];
    uint32_t loc[EPT];
    const size_t wbase = blk0 + (size_t)w * 64 * EPT + lane;
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t idx = wbase + (size_t)e * 64;
        bool valid = idx < n;
        key[e] = valid ? keys_in[idx] : K(0);
        val[e] = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
    }

#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t idx = wbase + (size_t)e * 64;
        bool valid = idx < n;
        uint32_t d = digit_of(key[e], shift, mask);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < 8; ++bit) {
            if (bit < width) {
                bool set = (d >> bit) & 1u;
                uint64_t bal = __ballot(set);
                peers &= set ? bal : ~bal;
            }
        }
        uint32_t rank = mask_rank(peers);
        uint32_t base = 0;
        if (valid) base = s_wh[w][d];
        if (valid && rank == 0) s_wh[w][d] = base + (uint32_t)__popcll(peers);
        loc[e] = base + rank;
    }
    __syncthreads();

    // Per-digit block start, then per-wave bases.
    {
        uint32_t cnt = 0;
        if (t < (uint32_t)RADIX) {
#pragma unroll
            for (uint32_t k = 0; k < NW; ++k) cnt += s_wh[k][t];
        }
        uint32_t all;
        uint32_t start = block_excl_scan_u32<NT>(cnt, s_tmp, &all);
        if (t < (uint32_t)RADIX) {
            uint32_t run = start;
#pragma unroll
            for (uint32_t k = 0; k < NW; ++k) {
                uint32_t c = s_wh[k][t];
                s_wh[k][t] = run;
                run += c;
            }
            s_gofs[t] -= start;
        }
    }
    __syncthreads();

#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        size_t idx = wbase + (size_t)e * 64;
        if (idx < n) {
            uint32_t d = digit_of(key[e], shift, mask);
            uint32_t pos = s_wh[w][d] + loc[e];
            s_keys[pos] = key[e];
            s_vals[pos] = val[e];
        }
    }
    __syncthreads();

#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        uint32_t s = t + e * NT;
        if (s < live) {
            K k = s_keys[s];
            uint32_t d = digit_of(k, shift, mask);
            size_t dst = (size_t)(s_gofs[d] + s);
            if (keys_out) keys_out[dst] = k;
            vals_out[dst] = s_vals[s];
        }
    }
}

template <typename K, int RADIX>
hipError_t pass_impl(hipStream_t s, const K* keys_in, const uint32_t* vals_in, K* keys_out,
                     uint32_t* vals_out, uint32_t n, const uint32_t* d_n, uint32_t shift,
                     uint32_t width, uint32_t* hist, uint32_t max_blocks, uint32_t* totals) {
    uint32_t grid = div_up(n, KPB);
    if (grid == 0) return hipSuccess;
    uint32_t mask = (1u << width) - 1u;
    k_upsweep<K, RADIX><<<grid, NT, 0, s>>>(keys_in, n, d_n, shift, mask, hist, max_blocks);
    k_digscan<RADIX><<<1u << width, NT, 0, s>>>(hist, max_blocks, n, d_n, totals);
    k_scatter<K, RADIX><<<grid, NT, 0, s>>>(keys_in, vals_in, keys_out, vals_out, n, d_n, shift,
                                            width, hist, max_blocks, totals);
    return hipGetLastError();
}

template <typename K>
hipError_t pass_dispatch(hipStream_t s, const K* keys_in, const uint32_t* vals_in, K* keys_out,
                         uint32_t* vals_out, uint32_t n, const uint32_t* d_n, uint32_t shift,
                         uint32_t width, uint32_t digit_bits, uint32_t* hist, uint32_t max_blocks,
                         uint32_t* totals) {
    // The LDS histogram is sized by the configured digit width; a narrower
    // final digit runs in the same instantiation.
    if (digit_bits <= 4)
        return pass_impl<K, 16>(s, keys_in, vals_in, keys_out, vals_out, n, d_n, shift, width,
                                hist, max_blocks, totals);
    if (digit_bits <= 6)
        return pass_impl<K, 64>(s, keys_in, vals_in, keys_out, vals_out, n, d_n, shift, width,
                                hist, max_blocks, totals);
    return pass_impl<K, 256>(s, keys_in, vals_in, keys_out, vals_out, n, d_n, shift, width, hist,
                             max_blocks, totals);
}

}  // namespace

uint32_t radix_max_blocks(uint32_t max_elems) { return div_up(max_elems > 0 ? max_elems : 1, KPB); }

hipError_t launch_radix_pass_u32(hipStream_t s, const uint32_t* keys_in, const uint32_t* vals_in,
                                 uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                 const uint32_t* d_n, uint32_t shift, uint32_t width,
                                 uint32_t digit_bits, uint32_t* hist, uint32_t max_blocks,
                                 uint32_t* totals) {
    return pass_dispatch<uint32_t>(s, keys_in, vals_in, keys_out, vals_out, n, d_n, shift, width,
                                   digit_bits, hist, max_blocks, totals);
}

hipError_t launch_radix_pass_u64(hipStream_t s, const uint64_t* keys_in, const uint32_t* vals_in,
                                 uint64_t* keys_out, uint32_t* vals_out, uint32_t n,
                                 const uint32_t* d_n, uint32_t shift, uint32_t width,
                                 uint32_t digit_bits, uint32_t* hist, uint32_t max_blocks,
                                 uint32_t* totals) {
    return pass_dispatch<uint64_t>(s, keys_in, vals_in, keys_out, vals_out, n, d_n, shift, width,
                                   digit_bits, hist, max_blocks, totals);
}

}  // namespace sfk
