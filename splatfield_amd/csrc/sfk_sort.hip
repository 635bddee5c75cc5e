// Stable LSD radix argsort for gfx950 (the contract of
// crates/sort/src/sort.rs:302-381, re-designed for 64-lane wavefronts).
//
// One pass over a `width`-bit digit (<= 8) is three launches:
//   upsweep   per-block digit histogram (LDS, per-wave sub-histograms),
//             written digit-major per segment: hist[(s*RADIX + d)*stride + b];
//   digscan   one workgroup per (segment, digit): exclusive scan of its row
//             over the segment's blocks, row total into totals[s*RADIX + d];
//   scatter   per block: stable rank of every key inside the block with
//             64-lane ballot match-any (one 64-bit ballot per digit bit; no
//             32-lane assumption as in sort.rs:57-87), LDS staging in sorted
//             order, then coalesced runs to segment base + digit prefix +
//             block prefix.
// Block = 4096 keys (2048 for small sorts) over kSortThreads threads; wave w
// owns keys [w*64*EPT, (w+1)*64*EPT) of its block, item i lane l ->
// w*64*EPT + i*64 + l, so ranks follow input order.  512 threads of 8 keys
// beat 256 of 16 (measured): each wave's ranking walks a chain of dependent
// LDS counter updates, one per key, and shorter chains over more waves
// overlap better.
//
// Segments: nseg independent sorts of seg_len keys each, laid out back to
// back (segment s at [s*seg_len, (s+1)*seg_len)), sorted in the same launches
// (the batched multi-view depth sort).  With one segment the element count
// may come from device memory (min(*d_n, n)): blocks past the live count exit
// at once, so a capacity-sized grid needs no host sync.  Identity values
// (vals_in == nullptr) are indices within the segment; with 8-byte values
// an identity value is (index, payload[index]), so per-element data rides
// along and comes out in sorted order without a gather.
#include "sfk_device.h"
#include "sfk_internal.h"

#include <algorithm>
#include <cstring>

namespace sfk {

namespace {

constexpr uint32_t NT = kSortThreads;
constexpr uint32_t NW = NT / 64;
constexpr uint32_t kEptBig = 4096 / NT, kEptSmall = 2048 / NT;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// The key as pass 0 sees it when rebasing (SortArgs::rebase_mode): fixed
// per segment, applied with selects only, after the block's loads are all
// issued (a branch per key on a loaded value makes the compiler wait for each
// load before issuing the next).
struct KeyBase {
    uint32_t on, base;
};

__device__ __forceinline__ KeyBase key_base(const SortArgs& a, uint32_t seg) {
    KeyBase r{0u, 0u};
    if (a.rebase_mode) {
        r.on = 1u;
        r.base = a.rebase_mode == 2 ? a.rebase_bases[seg] : a.rebase_base;
    }
    return r;
}

template <typename K>
__device__ __forceinline__ K rebased(K key, KeyBase kb) {
    if constexpr (sizeof(K) == 4) {
        const uint32_t d = key >= kb.base ? key - kb.base : 0u;
        const uint32_t c = d < 0xfffffeu ? d : 0xfffffeu;
        const uint32_t r = key == 0xffffffffu ? key : c;
        return kb.on ? r : key;
    } else {
        return key;
    }
}

// Which segment and block-within-segment this workgroup handles, and the
// segment's key range; false if the block has no live keys.  Blocks (chunks
// of kpb keys, numbered segment-major) go to workgroups through xcd_remap, so
// each XCD takes one contiguous range of them: the runs that consecutive
// blocks scatter into a digit's output range, and their entries of a
// histogram row, are then adjacent within one L2 and merge there into whole
// cache lines, where round-robin blocks would each write back partial lines.
struct SegPos {
    uint32_t s, bl, nb, seg_n;
    size_t seg_base;
};

__device__ __forceinline__ bool seg_pos(const SortArgs& a, uint32_t b, SegPos& p) {
    if (a.nseg <= 1) {
        uint32_t n = a.n_max;
        if (a.d_n) {
            uint32_t m = *a.d_n;
            n = m < n ? m : n;
        }
        p.nb = (n + a.kpb - 1) / a.kpb;
        if (b >= p.nb) return false;
        p.s = 0;
        p.bl = xcd_remap(b, p.nb);
        p.seg_n = n;
        p.seg_base = 0;
    } else {
        const uint32_t bps = (a.seg_len + a.kpb - 1) / a.kpb;
        const uint32_t all = bps * a.nseg;
        if (b >= all) return false;
        const uint32_t c = xcd_remap(b, all);
        p.s = c / bps;
        p.bl = c % bps;
        p.nb = bps;
        p.seg_n = a.seg_len;
        p.seg_base = (size_t)p.s * a.seg_len;
    }
    return true;
}

template <typename K, int RADIX, uint32_t EPT>
__device__ __forceinline__ void upsweep_block(const SortArgs& a, const SegPos& p) {
    constexpr uint32_t KPB = NT * EPT;
    __shared__ uint32_t h[NW][RADIX];
    const K* keys = static_cast<const K*>(a.keys_in);
    const uint32_t w = threadIdx.x / 64;
    const uint32_t mask = (1u << a.width) - 1u;
    for (uint32_t i = threadIdx.x; i < NW * RADIX; i += NT) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t loc0 = p.bl * KPB + threadIdx.x;
    // All loads first, then the LDS counts, so the loads overlap; a block
    // fully inside its segment loads without per-element guards.
    K kr[EPT];
    const K* kp = keys + p.seg_base + loc0;
    const bool full = p.bl * KPB + KPB <= p.seg_n;
    if (full) {
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) kr[e] = kp[e * NT];
    } else {
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) kr[e] = loc0 + e * NT < p.seg_n ? kp[e * NT] : K(0);
    }
    const KeyBase kb = key_base(a, p.s);
    uint32_t dg[EPT];
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e)
        dg[e] = full || loc0 + e * NT < p.seg_n ? digit_of(rebased(kr[e], kb), a.shift, mask) : ~0u;
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e)
        if (dg[e] != ~0u) atomicAdd(&h[w][dg[e]], 1u);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < RADIX; d += NT) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t k = 0; k < NW; ++k) t += h[k][d];
        a.hist[((size_t)p.s * RADIX + d) * a.stride + p.bl] = t;
    }
}

// LOOP: blocks b = blockIdx.x, + gridDim.x, ... while live.  A grid sized to
// a capacity (the element count read from device memory) is bounded on the
// host instead of launching a workgroup per capacity block that would only
// load the count and exit; gridDim.x is then a multiple of 8, so block b still
// runs on XCD b % 8 (see seg_pos).  Without LOOP, one block per workgroup.
template <typename K, int RADIX, uint32_t EPT, bool LOOP>
__global__ __launch_bounds__(NT) void k_upsweep(SortArgs a) {
    SegPos p;
    if constexpr (LOOP) {
        for (uint32_t b = blockIdx.x; seg_pos(a, b, p); b += gridDim.x) {
            upsweep_block<K, RADIX, EPT>(a, p);
            __syncthreads();  // LDS reused by the next block
        }
    } else {
        if (seg_pos(a, blockIdx.x, p)) upsweep_block<K, RADIX, EPT>(a, p);
    }
}

// One workgroup per (segment, digit): exclusive scan of its row in place.
template <int RADIX>
__global__ __launch_bounds__(NT) void k_digscan(SortArgs a) {
    __shared__ uint32_t tmp[NW + 1];
    const uint32_t nbins = 1u << a.width;
    const uint32_t s = blockIdx.x / nbins, d = blockIdx.x % nbins;
    uint32_t nb;
    if (a.nseg <= 1) {
        uint32_t n = a.n_max;
        if (a.d_n) {
            uint32_t m = *a.d_n;
            n = m < n ? m : n;
        }
        nb = (n + a.kpb - 1) / a.kpb;
    } else {
        nb = (a.seg_len + a.kpb - 1) / a.kpb;
    }
    uint32_t* row = a.hist + ((size_t)s * RADIX + d) * a.stride;
    uint32_t carry = 0;
    constexpr uint32_t EPT = 16;
    for (uint32_t c0 = 0; c0 < nb; c0 += NT * EPT) {
        uint32_t v[EPT];
        uint32_t sum = 0;
        uint32_t first = c0 + threadIdx.x * EPT;
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            uint32_t idx = first + e;
            v[e] = idx < nb ? row[idx] : 0u;
            sum += v[e];
        }
        uint32_t total;
        uint32_t ex = block_excl_scan_u32<NT>(sum, tmp, &total);
        uint32_t run = carry + ex;
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            uint32_t idx = first + e;
            if (idx < nb) row[idx] = run;
            run += v[e];
        }
        carry += total;
    }
    if (threadIdx.x == 0) a.totals[(size_t)s * RADIX + d] = carry;
}

// The same scans with a wavefront per (segment, digit) row, for rows of at
// most kRowScanMax blocks: coalesced loads of 64 consecutive counts, a DPP
// scan each, no barrier.  (A workgroup per row of a few hundred counts is
// mostly idle threads and two barriers.)
constexpr uint32_t kRowScanMax = 512;
__device__ __forceinline__ uint32_t scan_blocks(const SortArgs& a) {
    if (a.nseg <= 1) {
        uint32_t n = a.n_max;
        if (a.d_n) {
            const uint32_t m = *a.d_n;
            n = m < n ? m : n;
        }
        return (n + a.kpb - 1) / a.kpb;
    }
    return (a.seg_len + a.kpb - 1) / a.kpb;
}
template <int RADIX>
__global__ __launch_bounds__(NT) void k_digscan_rows(SortArgs a, uint32_t rows) {
    const uint32_t q = blockIdx.x * NW + threadIdx.x / 64;
    if (q >= rows) return;
    // row q = (segment, digit) of the pass's 2^width digits
    const size_t r = (size_t)(q >> a.width) * RADIX + (q & ((1u << a.width) - 1u));
    const uint32_t lane = lane_id();
    const uint32_t nb = scan_blocks(a);
    uint32_t* row = a.hist + r * a.stride;
    constexpr uint32_t C = kRowScanMax / 64;
    uint32_t v[C];
#pragma unroll
    for (uint32_t e = 0; e < C; ++e) {
        const uint32_t idx = e * 64 + lane;
        v[e] = idx < nb ? row[idx] : 0u;
    }
    uint32_t carry = 0;
#pragma unroll
    for (uint32_t e = 0; e < C; ++e) {
        const uint32_t idx = e * 64 + lane;
        const uint32_t inc = wave_incl_scan_u32(v[e]);
        if (idx < nb) row[idx] = carry + inc - v[e];
        carry += __builtin_amdgcn_readlane(inc, 63);
    }
    if (lane == 0) a.totals[r] = carry;
}

template <typename V>
__device__ __forceinline__ V identity_val(uint32_t li, const uint32_t* payload, size_t gi);
template <>
__device__ __forceinline__ uint32_t identity_val<uint32_t>(uint32_t li, const uint32_t*, size_t) {
    return li;
}
template <>
__device__ __forceinline__ uint2 identity_val<uint2>(uint32_t li, const uint32_t* payload,
                                                     size_t gi) {
    return make_uint2(li, payload[gi]);
}

template <typename K, typename V, int RADIX, uint32_t EPT>
__device__ __forceinline__ void scatter_block(const SortArgs& a, const SegPos& p) {
    constexpr uint32_t KPB = NT * EPT;
    // The digit width is log2(RADIX) (RADIX = 16 also serves widths below 4:
    // the digit's upper bits are then zero in every lane).
    constexpr uint32_t kDigitBits = RADIX >= 256 ? 8 : RADIX >= 128 ? 7 : RADIX >= 64 ? 6 : RADIX >= 32 ? 5 : 4;
    // Keys, then values, staged in sorted order through the same bytes: a
    // block holds the larger of the two instead of both (more workgroups per
    // CU for 4096-key blocks of u64 keys or 8-byte values).
    constexpr size_t kStage = KPB * (sizeof(K) > sizeof(V) ? sizeof(K) : sizeof(V));
    __shared__ __align__(16) unsigned char s_stage[kStage];
    K* s_keys = reinterpret_cast<K*>(s_stage);
    V* s_vals = reinterpret_cast<V*>(s_stage);
    __shared__ uint32_t s_wh[NW][RADIX];  // per-wave digit counters, then bases
    __shared__ uint32_t s_gofs[RADIX];    // global position minus local position
    __shared__ uint32_t s_tmp[NW + 1];

    const K* keys_in = static_cast<const K*>(a.keys_in);
    K* keys_out = static_cast<K*>(a.keys_out);
    const uint32_t t = threadIdx.x, w = t / 64, lane = lane_id();
    const uint32_t width = a.width, shift = a.shift;
    const uint32_t mask = (1u << width) - 1u;
    const uint32_t nbins = 1u << width;
    const uint32_t blk0 = p.bl * KPB;  // within the segment
    const uint32_t live = (p.seg_n - blk0) < KPB ? (p.seg_n - blk0) : KPB;

    for (uint32_t i = t; i < NW * RADIX; i += NT) (&s_wh[0][0])[i] = 0;

    K key[EPT];
    V val[EPT];
    const V* vals_in = static_cast<const V*>(a.vals_in);
    V* vals_out = static_cast<V*>(a.vals_out);
    uint32_t loc[EPT];
    const uint32_t wbase = blk0 + w * 64 * EPT + lane;
    const size_t gb = p.seg_base;
    // Every load of the block is issued before any is used (see KeyBase).
    if (blk0 + KPB <= p.seg_n) {
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) key[e] = keys_in[gb + wbase + e * 64];
        if (vals_in) {
#pragma unroll
            for (uint32_t e = 0; e < EPT; ++e) val[e] = vals_in[gb + wbase + e * 64];
        } else {
#pragma unroll
            for (uint32_t e = 0; e < EPT; ++e)
                val[e] = identity_val<V>(wbase + e * 64, a.payload, gb + wbase + e * 64);
        }
    } else {
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) {
            const uint32_t li = wbase + e * 64;
            const bool valid = li < p.seg_n;
            key[e] = valid ? keys_in[gb + li] : K(0);
            val[e] = valid ? (vals_in ? vals_in[gb + li] : identity_val<V>(li, a.payload, gb + li))
                           : V{};
        }
    }
    // Segment-global digit base (exclusive scan of the segment's digit
    // totals) while the block's keys and values are in flight.
    {
        uint32_t tot = t < nbins ? a.totals[(size_t)p.s * RADIX + t] : 0u;
        uint32_t all;
        uint32_t ex = block_excl_scan_u32<NT>(tot, s_tmp, &all);
        if (t < (uint32_t)RADIX)
            s_gofs[t] = t < nbins ? (uint32_t)p.seg_base + ex +
                                        a.hist[((size_t)p.s * RADIX + t) * a.stride + p.bl]
                                  : 0u;
    }
    // (the scan's barriers also order the s_wh clear)
    {
        const KeyBase kb = key_base(a, p.s);
#pragma unroll
        for (uint32_t e = 0; e < EPT; ++e) key[e] = rebased(key[e], kb);
    }

#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        bool valid = wbase + e * 64 < p.seg_n;
        uint32_t d = digit_of(key[e], shift, mask);
        const uint64_t peers = match_peers<kDigitBits>(d, valid);
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

    // loc[e] becomes the key's place in the block's sorted order.
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        if (wbase + e * 64 < p.seg_n) {
            uint32_t d = digit_of(key[e], shift, mask);
            loc[e] += s_wh[w][d];
            s_keys[loc[e]] = key[e];
        }
    }
    __syncthreads();

    uint32_t dst[EPT];  // segment-global position of sorted slot t + e * NT
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        uint32_t s = t + e * NT;
        if (s < live) {
            K k = s_keys[s];
            uint32_t d = digit_of(k, shift, mask);
            dst[e] = s_gofs[d] + s;
            if (keys_out) keys_out[dst[e]] = k;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e)
        if (wbase + e * 64 < p.seg_n) s_vals[loc[e]] = val[e];
    __syncthreads();
#pragma unroll
    for (uint32_t e = 0; e < EPT; ++e) {
        uint32_t s = t + e * NT;
        if (s < live) vals_out[dst[e]] = s_vals[s];
    }
}

template <typename K, typename V, int RADIX, uint32_t EPT, bool LOOP>
__global__ __launch_bounds__(NT) void k_scatter(SortArgs a) {
    SegPos p;
    if constexpr (LOOP) {
        for (uint32_t b = blockIdx.x; seg_pos(a, b, p); b += gridDim.x) {
            scatter_block<K, V, RADIX, EPT>(a, p);
            __syncthreads();  // LDS reused by the next block
        }
    } else {
        if (seg_pos(a, blockIdx.x, p)) scatter_block<K, V, RADIX, EPT>(a, p);
    }
}

// Workgroups of a pass whose element count is read from device memory.
constexpr uint32_t kMaxLiveGrid = 2048;
static_assert(kMaxLiveGrid % 8 == 0, "keeps block b on XCD b % 8");

template <typename K, typename V, int RADIX, uint32_t EPT>
hipError_t pass_impl(hipStream_t st, const SortArgs& a) {
    constexpr uint32_t KPB = NT * EPT;
    uint32_t grid = a.nseg <= 1 ? div_up(a.n_max, KPB) : a.nseg * div_up(a.seg_len, KPB);
    if (grid == 0) return hipSuccess;
    // Rows of the digit scan: at most `grid` blocks each (fewer when the
    // count read from device memory is smaller).
    const uint32_t rows = (a.nseg <= 1 ? 1u : a.nseg) << a.width;
    const uint32_t row_blocks = a.nseg <= 1 ? grid : div_up(a.seg_len, KPB);
    auto digscan = [&]() {
        if (row_blocks <= kRowScanMax)
            k_digscan_rows<RADIX><<<div_up(rows, NW), NT, 0, st>>>(a, rows);
        else
            k_digscan<RADIX><<<rows, NT, 0, st>>>(a);
    };
    if (a.nseg <= 1 && a.d_n && grid > kMaxLiveGrid) {
        k_upsweep<K, RADIX, EPT, true><<<kMaxLiveGrid, NT, 0, st>>>(a);
        digscan();
        k_scatter<K, V, RADIX, EPT, true><<<kMaxLiveGrid, NT, 0, st>>>(a);
    } else {
        k_upsweep<K, RADIX, EPT, false><<<grid, NT, 0, st>>>(a);
        digscan();
        k_scatter<K, V, RADIX, EPT, false><<<grid, NT, 0, st>>>(a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ bucket finish
//
// A u64 sort of `bits` >= 24 bits whose keys spread over their top 16 bits:
// two stable passes on those 16 bits (the three-launch passes above) leave
// the keys grouped by their top 16 bits, each group (bucket) in input order;
// sorting every bucket on its own by the remaining bits, stably, completes
// the stable sort by all `bits`.  k_bucket_first marks where each bucket
// starts, k_bucket_sizes turns that into sizes and their maximum (the host
// reads it: every bucket must fit kBucketCap, else the plain passes run
// instead), and k_bucket_sort sorts each bucket with one wavefront on chip:
// keys and 16-bit slots through LDS, LSD passes of at most 8 bits ranked by
// 64-lane ballot match, values gathered once by slot.

// first[t] / last[t] = indices of the first and last key of bucket t in the
// grouped keys (written only for nonempty buckets, each by the one thread
// at that boundary; the caller fills first with 0xffffffff).
// A wavefront takes 64 * kBoundsKpt consecutive keys, step j lane l key
// 64j + l, and also loads the 64 keys on either side of its range (steps
// -1 and kBoundsKpt; neighbouring wavefronts read them too, so they mostly
// hit in L2).  Every load is issued before any compare (indices clamped, no
// branch around a load).  A key's neighbours come from the adjacent lanes,
// across steps from lane 63 of the step before / lane 0 of the step after.
constexpr uint32_t kBoundsKpt = 16;
__global__ __launch_bounds__(256) void k_bucket_bounds(const uint64_t* __restrict__ keys, uint32_t n,
                                                       uint32_t shift, uint32_t* __restrict__ first,
                                                       uint32_t* __restrict__ last) {
    constexpr uint32_t kSteps = kBoundsKpt + 2;
    const uint32_t lane = lane_id();
    const int64_t wbase = ((int64_t)blockIdx.x * 4 + threadIdx.x / 64) * (64 * kBoundsKpt);
    if (wbase >= (int64_t)n) return;
    const int64_t top = (int64_t)n - 1;
    constexpr uint32_t kNone = 0xffffffffu;
    uint32_t t[kSteps];  // t[j + 1] = bucket of key wbase + 64j + lane
#pragma unroll
    for (uint32_t s = 0; s < kSteps; ++s) {
        const int64_t i = wbase + 64 * ((int64_t)s - 1) + lane;
        const uint64_t k = keys[i < 0 ? 0 : (i < top ? i : top)];
        t[s] = i >= 0 && i <= top ? (uint32_t)(k >> shift) & 0xffffu : kNone;
    }
#pragma unroll
    for (uint32_t s = 1; s <= kBoundsKpt; ++s) {
        uint32_t tp = (uint32_t)__shfl_up((int)t[s], 1, 64);
        uint32_t tn = (uint32_t)__shfl_down((int)t[s], 1, 64);
        const uint32_t before = (uint32_t)__builtin_amdgcn_readlane((int)t[s - 1], 63);
        const uint32_t after = (uint32_t)__builtin_amdgcn_readlane((int)t[s + 1], 0);
        if (lane == 0) tp = before;
        if (lane == 63) tn = after;
        const int64_t i = wbase + 64 * ((int64_t)s - 1) + lane;
        if (i <= top) {
            if (tp != t[s]) first[t[s]] = (uint32_t)i;
            if (tn != t[s]) last[t[s]] = (uint32_t)i;
        }
    }
}

// size[t] of every bucket (empty: 0) and, per workgroup of 256 buckets, the
// largest (wg_max[blockIdx.x]; the host takes the maximum of the 256).
constexpr uint32_t kBucketSizeGroups = 65536 / 256;
__global__ __launch_bounds__(256) void k_bucket_sizes(const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ last,
                                                      uint32_t* __restrict__ size,
                                                      uint32_t* __restrict__ wg_max) {
    __shared__ uint32_t s_max[4];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t f = first[t];
    const uint32_t sz = f == 0xffffffffu ? 0u : last[t] - f + 1u;
    size[t] = sz;
    uint32_t mx = sz;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
    if (lane_id() == 0) s_max[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) wg_max[blockIdx.x] = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

// One wavefront per bucket (four per workgroup): the bucket's keys (u64,
// at most kBucketCap) sorted stably by bits [0, lbits); keys_in / vals_in the
// grouped arrays, keys_out / vals_out the result (keys_out may be null).
//  * lbits <= 16: LSD passes of width[p] bits over all of them.
//  * lbits > 16: rank by the top kPfxBits of them through a presence
//    bitmap in LDS (one bit per prefix value): a key's rank is the number of
//    set bits below its prefix, a popcount over its 128-bit group plus the
//    group's prefix count.  When every prefix is distinct (about a third of
//    random buckets of ~256 keys), those ranks are the final places.  Else a
//    count per distinct prefix gives each of its keys a place (in arbitrary
//    order within the prefix); when no prefix is shared by more than two
//    keys, one compare and swap per shared pair on (remaining bits, input
//    position) restores the stable order.  A bucket where three or more
//    keys share a prefix (ties, clustered keys) is sorted from its input
//    order by the full LSD passes instead.
struct BucketArgs {
    const uint64_t* keys_in;
    const uint32_t* vals_in;
    uint64_t* keys_out;
    uint32_t* vals_out;
    const uint32_t* first;
    const uint32_t* size;
    uint32_t lbits;
    uint32_t passes;
    uint32_t width[7];
};
constexpr uint32_t kBucketEpt = 6;
constexpr uint32_t kBucketCap = 64 * kBucketEpt;
static_assert(kBucketCap <= 65536, "16-bit slots");

// One stable LSD pass of the wave's items (item e of lane l is element
// e * 64 + l, e < ept) on the `width`-bit digit at `sh`: ranks by 64-lane
// ballot match, digit starts from a 256-entry count, keys and slots through
// LDS back into the same item order.
__device__ __forceinline__ void bucket_pass(uint64_t (&key)[kBucketEpt], uint32_t (&pos)[kBucketEpt],
                                            const bool (&valid)[kBucketEpt], uint32_t ept,
                                            uint32_t sh, uint32_t width, uint32_t* cnt,
                                            uint64_t* s_key, uint16_t* s_idx) {
    const uint32_t lane = lane_id();
    const uint32_t mask = (1u << width) - 1u;
    uint32_t loc[kBucketEpt], dg[kBucketEpt];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) cnt[lane + 64 * k] = 0u;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        if (e >= ept) break;
        const uint32_t d = (uint32_t)(key[e] >> sh) & mask;
        dg[e] = d;
        const uint64_t peers = match_peers<8>(d, valid[e]);
        const uint32_t rank = mask_rank(peers);
        uint32_t base = 0;
        if (valid[e]) base = cnt[d];
        if (valid[e] && rank == 0) cnt[d] = base + (uint32_t)__popcll(peers);
        loc[e] = base + rank;
    }
    __builtin_amdgcn_wave_barrier();
    {
        // Exclusive scan of the 256 counts: lane l holds digits 4l .. 4l + 3.
        uint32_t c[4], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            c[k] = cnt[4 * lane + k];
            sum += c[k];
        }
        uint32_t run = wave_incl_scan_u32(sum) - sum;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            cnt[4 * lane + k] = run;
            run += c[k];
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        if (e >= ept) break;
        if (valid[e]) {
            const uint32_t q = cnt[dg[e]] + loc[e];
            s_key[q] = key[e];
            s_idx[q] = (uint16_t)pos[e];
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        if (e >= ept) break;
        const uint32_t i = e * 64 + lane;
        if (valid[e]) {
            key[e] = s_key[i];
            pos[e] = s_idx[i];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// Wave-synchronous LDS hand-off: the wave's earlier LDS accesses are seen
// by its later ones (LDS serves one wave's instructions in order; the fences
// keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kPfxBits = 15;
constexpr uint32_t kMapWords = (1u << kPfxBits) / 32;  // presence bitmap
constexpr uint32_t kMapGroups = kMapWords / 4;         // 128-bit groups
static_assert(kMapGroups == 4 * 64, "four bitmap groups per lane");
// Per-wave LDS region, used twice:
//   ranking: bitmap [0, kMapWords*4), group prefix counts (u16) after it;
//   placing: prefix-class counts (u32) [0, kCntBytes), then the values in
//            sorted order (u32) in the same bytes; sorted keys (u64) at
//            kKeyOff, sorted input positions (u16) at kIdxOff; the full
//            passes' digit counts (256 u32) reuse [0, 1024).
constexpr uint32_t kCntBytes = kBucketCap * 4;
constexpr uint32_t kKeyOff = kCntBytes;
constexpr uint32_t kIdxOff = kKeyOff + kBucketCap * 8;
constexpr uint32_t kRegionBytes =
    std::max(kIdxOff + kBucketCap * 2, kMapWords * 4 + kMapGroups * 2);
static_assert(kRegionBytes % 16 == 0 && kCntBytes >= 1024, "region layout");

__global__ __launch_bounds__(256) void k_bucket_sort(BucketArgs a) {
    __shared__ __align__(16) unsigned char s_mem[4][kRegionBytes];
    const uint32_t w = threadIdx.x / 64, lane = lane_id();
    const uint32_t b = blockIdx.x * 4u + w;
    const uint32_t m = a.size[b];
    if (m == 0 || m > kBucketCap) return;  // (sizes above the cap never reach here)
    unsigned char* reg = s_mem[w];
    uint32_t* map = reinterpret_cast<uint32_t*>(reg);
    uint16_t* gpre = reinterpret_cast<uint16_t*>(reg + kMapWords * 4);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(reg);
    uint32_t* s_val = reinterpret_cast<uint32_t*>(reg);
    uint64_t* s_key = reinterpret_cast<uint64_t*>(reg + kKeyOff);
    uint16_t* s_idx = reinterpret_cast<uint16_t*>(reg + kIdxOff);
    const size_t start = a.first[b];
    const uint32_t ept = (m + 63) / 64;  // wave-uniform, 1..kBucketEpt
    uint64_t key[kBucketEpt];
    uint32_t val[kBucketEpt];
    bool valid[kBucketEpt];
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        const uint32_t i = e * 64 + lane;
        valid[e] = e < ept && i < m;
        key[e] = valid[e] ? a.keys_in[start + i] : 0ull;
        val[e] = valid[e] ? a.vals_in[start + i] : 0u;
    }
    bool need_full = a.lbits <= 16;
    if (!need_full) {
        const uint32_t hs = a.lbits - kPfxBits;
        uint32_t pf[kBucketEpt], q[kBucketEpt], sub[kBucketEpt];
        // Presence bitmap of the prefixes.
#pragma unroll
        for (uint32_t j = 0; j < kMapWords / 256; ++j)
            reinterpret_cast<uint4*>(map)[lane + 64 * j] = make_uint4(0u, 0u, 0u, 0u);
        wave_lds_sync();
#pragma unroll
        for (uint32_t e = 0; e < kBucketEpt; ++e) {
            pf[e] = (uint32_t)(key[e] >> hs) & ((1u << kPfxBits) - 1u);
            if (valid[e]) atomicOr(&map[pf[e] >> 5], 1u << (pf[e] & 31u));
        }
        wave_lds_sync();
        // Set bits per 128-bit group (lane l: groups l, l+64, l+128, l+192),
        // then the exclusive count of set bits before every group; two groups
        // per 16-bit field of one scan (counts stay below 2^16).
        uint32_t c[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint4 v = reinterpret_cast<const uint4*>(map)[lane + 64 * j];
            c[j] = __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
        }
        const uint32_t x0 = c[0] | (c[1] << 16), x1 = c[2] | (c[3] << 16);
        const uint32_t i0 = wave_incl_scan_u32(x0), i1 = wave_incl_scan_u32(x1);
        const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)i0, 63);
        const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)i1, 63);
        const uint32_t r0 = t0 & 0xffffu, r1 = t0 >> 16, r2 = t1 & 0xffffu, r3 = t1 >> 16;
        const uint32_t e0 = i0 - x0, e1 = i1 - x1;
        gpre[lane] = (uint16_t)(e0 & 0xffffu);
        gpre[lane + 64] = (uint16_t)(r0 + (e0 >> 16));
        gpre[lane + 128] = (uint16_t)(r0 + r1 + (e1 & 0xffffu));
        gpre[lane + 192] = (uint16_t)(r0 + r1 + r2 + (e1 >> 16));
        const uint32_t distinct = r0 + r1 + r2 + r3;
        wave_lds_sync();
        // Rank of each key's prefix among the distinct prefixes.
#pragma unroll
        for (uint32_t e = 0; e < kBucketEpt; ++e) {
            if (e >= ept) break;
            const uint32_t g = pf[e] >> 7, sb = pf[e] & 127u, k = sb >> 5;
            const uint4 v = reinterpret_cast<const uint4*>(map)[g];
            const uint32_t wk = k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
            uint32_t r = gpre[g] + __popc(wk & ((1u << (sb & 31u)) - 1u));
            r += k > 0 ? __popc(v.x) : 0u;
            r += k > 1 ? __popc(v.y) : 0u;
            r += k > 2 ? __popc(v.z) : 0u;
            q[e] = r;
            sub[e] = 0;
        }
        wave_lds_sync();
        const bool shared = distinct != m;  // some prefix held by two or more keys
        if (shared) {
            for (uint32_t i = lane; i < distinct; i += 64) cnt[i] = 0u;
            wave_lds_sync();
            bool many = false;
#pragma unroll
            for (uint32_t e = 0; e < kBucketEpt; ++e) {
                if (valid[e]) {
                    sub[e] = atomicAdd(&cnt[q[e]], 1u);
                    many |= sub[e] >= 2;
                }
            }
            wave_lds_sync();
            if (__ballot(many)) {
                need_full = true;
            } else {
                // Exclusive scan of the class counts: lane l, classes 6l .. 6l+5.
                constexpr uint32_t kPer = kBucketCap / 64;
                uint32_t cc[kPer], sum = 0;
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t i = kPer * lane + k;
                    cc[k] = i < distinct ? cnt[i] : 0u;
                    sum += cc[k];
                }
                uint32_t run = wave_incl_scan_u32(sum) - sum;
                wave_lds_sync();
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t i = kPer * lane + k;
                    if (i < distinct) cnt[i] = run;
                    run += cc[k];
                }
                wave_lds_sync();
#pragma unroll
                for (uint32_t e = 0; e < kBucketEpt; ++e)
                    if (valid[e]) q[e] = cnt[q[e]] + sub[e];
            }
        }
        if (!need_full) {
            wave_lds_sync();  // (the class counts are dead: s_val takes their bytes)
#pragma unroll
            for (uint32_t e = 0; e < kBucketEpt; ++e) {
                if (valid[e]) {
                    s_key[q[e]] = key[e];
                    s_idx[q[e]] = (uint16_t)(e * 64 + lane);
                    s_val[q[e]] = val[e];
                }
            }
            wave_lds_sync();
            if (shared) {
                // The second key of a shared prefix (sub 1) sits right after
                // the first: order the pair by (remaining bits, position).
                const uint64_t low = (1ull << hs) - 1ull;
#pragma unroll
                for (uint32_t e = 0; e < kBucketEpt; ++e) {
                    if (valid[e] && sub[e] == 1u) {
                        const uint32_t qa = q[e] - 1u, pos = e * 64 + lane;
                        const uint64_t ka = s_key[qa];
                        const uint32_t pa = s_idx[qa];
                        const uint64_t lm = key[e] & low, la = ka & low;
                        if (lm < la || (lm == la && pos < pa)) {
                            const uint32_t va = s_val[qa];
                            s_key[qa] = key[e];
                            s_idx[qa] = (uint16_t)pos;
                            s_val[qa] = val[e];
                            s_key[q[e]] = ka;
                            s_idx[q[e]] = (uint16_t)pa;
                            s_val[q[e]] = va;
                        }
                    }
                }
                wave_lds_sync();
            }
        }
    }
    if (need_full) {
        uint32_t pos[kBucketEpt];
#pragma unroll
        for (uint32_t e = 0; e < kBucketEpt; ++e) pos[e] = e * 64 + lane;
        uint32_t sh = 0;
        for (uint32_t p = 0; p < a.passes; ++p) {
            bucket_pass(key, pos, valid, ept, sh, a.width[p], cnt, s_key, s_idx);
            sh += a.width[p];
        }
        // (bucket_pass leaves the sorted keys and positions in s_key / s_idx.)
    }
    // The full passes carry positions only: their values are gathered (all
    // gathers issued before any store).
    uint64_t ko[kBucketEpt];
    uint32_t v[kBucketEpt];
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        if (valid[e]) {
            const uint32_t i = e * 64 + lane;
            ko[e] = s_key[i];
            v[e] = need_full ? a.vals_in[start + s_idx[i]] : s_val[i];
        }
    }
#pragma unroll
    for (uint32_t e = 0; e < kBucketEpt; ++e) {
        if (valid[e]) {
            const size_t o = start + e * 64 + lane;
            if (a.keys_out) a.keys_out[o] = ko[e];
            a.vals_out[o] = v[e];
        }
    }
}

}  // namespace

hipError_t launch_bucket_split(hipStream_t st, const uint64_t* grouped_keys, uint32_t n,
                               uint32_t shift, uint32_t* first, uint32_t* last, uint32_t* size,
                               uint32_t* wg_max) {
    hipError_t e = launch_fill_u32(st, first, 0xffffffffu, 65536);
    if (e != hipSuccess) return e;
    k_bucket_bounds<<<div_up(n, 256u * kBoundsKpt), 256, 0, st>>>(grouped_keys, n, shift, first, last);
    k_bucket_sizes<<<kBucketSizeGroups, 256, 0, st>>>(first, last, size, wg_max);
    return hipGetLastError();
}

uint32_t bucket_size_groups() { return kBucketSizeGroups; }

uint32_t bucket_cap() { return kBucketCap; }

hipError_t launch_bucket_sort(hipStream_t st, const uint64_t* keys_in, const uint32_t* vals_in,
                              uint64_t* keys_out, uint32_t* vals_out, const uint32_t* first,
                              const uint32_t* size, uint32_t lbits) {
    BucketArgs a;
    memset(&a, 0, sizeof a);
    a.keys_in = keys_in;
    a.vals_in = vals_in;
    a.keys_out = keys_out;
    a.vals_out = vals_out;
    a.first = first;
    a.size = size;
    a.lbits = lbits;
    a.passes = (lbits + 7) / 8;
    if (a.passes == 0 || a.passes > 7) return hipErrorInvalidValue;
    for (uint32_t p = 0; p < a.passes; ++p)
        a.width[p] = lbits / a.passes + (p < lbits % a.passes ? 1u : 0u);
    k_bucket_sort<<<65536 / 4, 256, 0, st>>>(a);
    return hipGetLastError();
}

uint32_t radix_keys_per_block(uint64_t total) {
    return total <= kSortSmallMax ? NT * kEptSmall : NT * kEptBig;
}

// Histogram rows for any sort a scratch of max_elems keys can run.
uint32_t radix_max_blocks(uint32_t max_elems) {
    const uint64_t m = max_elems > 0 ? max_elems : 1;
    const uint32_t big = div_up(m, NT * kEptBig);
    const uint32_t small = div_up(std::min<uint64_t>(m, kSortSmallMax), NT * kEptSmall);
    return big > small ? big : small;
}

template <typename K, typename V, uint32_t EPT>
hipError_t pass_ept(hipStream_t st, const SortArgs& a, uint32_t digit_bits) {
    if (digit_bits <= 4) return pass_impl<K, V, 16, EPT>(st, a);
    if (digit_bits <= 5) return pass_impl<K, V, 32, EPT>(st, a);
    if (digit_bits <= 6) return pass_impl<K, V, 64, EPT>(st, a);
    if (digit_bits <= 7) return pass_impl<K, V, 128, EPT>(st, a);
    return pass_impl<K, V, 256, EPT>(st, a);
}

// Small sorts (up to kSortSmallMax keys) use 2048-key blocks: twice the
// workgroups keep more waves in flight where 4096-key blocks would leave one
// per CU (measured: 1024-key blocks are slower again).
template <typename K, typename V>
hipError_t pass_radix(hipStream_t st, const SortArgs& a, uint32_t digit_bits) {
    if (a.kpb == NT * kEptSmall) return pass_ept<K, V, kEptSmall>(st, a, digit_bits);
    return pass_ept<K, V, kEptBig>(st, a, digit_bits);
}

hipError_t launch_radix_pass(hipStream_t st, const SortArgs& a, uint32_t key_bytes,
                             uint32_t val_bytes, uint32_t digit_bits) {
    if (key_bytes == 4)
        return val_bytes == 8 ? pass_radix<uint32_t, uint2>(st, a, digit_bits)
                              : pass_radix<uint32_t, uint32_t>(st, a, digit_bits);
    return pass_radix<uint64_t, uint32_t>(st, a, digit_bits);
}

}  // namespace sfk
