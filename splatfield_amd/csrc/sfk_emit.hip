// Fused emission-offset scan and isect emission for gfx950: one pass over
// the depth ranks that does what scan.rs:215-255 (exclusive_scan_gather of
// tile counts in depth order) followed by raster.rs:12-48 (map_isects) do,
// with one gather of the per-splat counts and bboxes instead of three.
//
// Workgroups of 256 threads own 1024 consecutive ranks (4 per thread).  The
// workgroup prefix comes from a decoupled look-back over the workgroups
// before it: each publishes its aggregate, then its inclusive prefix, in one
// 64-bit status word (flag << 32 | u32 value).  Logical workgroup ids are
// taken from a ticket counter, so every workgroup a look-back waits on has
// already started and publishes its aggregate before it waits on anything:
// every wait ends.  Each wavefront then writes the isects of its 256 ranks,
// which form one contiguous output range, with consecutive lanes on
// consecutive slots (coalesced), finding each slot's rank by binary search
// over the wave's prefix sums in LDS.
#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

constexpr uint32_t kEmitThreads = 256;
constexpr uint32_t kEmitRpt = 4;  // ranks per thread
constexpr uint32_t kEmitRanks = kEmitThreads * kEmitRpt;
constexpr uint32_t kFlagAggregate = 1, kFlagPrefix = 2;

__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t* p, uint32_t flag, uint32_t v) {
    __hip_atomic_store(p, ((uint64_t)flag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct EmitArgs {
    const uint32_t* order;        // depth rank -> splat (in-view index when batched)
    const uint32_t* tile_counts;  // per splat (view-major when batched)
    const uint2* bbox;            // packed tile bbox per splat
    const uint2* ordered;         // or: (splat, compact bbox) per depth rank
    uint32_t* order_out;          // with `ordered`: depth order (optional)
    uint32_t n;                   // ranks
    uint32_t tiles_per_row;
    uint32_t max_isects;
    uint32_t* tile_ids;
    uint32_t* gaussian_ids;
    uint32_t* offsets;            // optional: per-rank exclusive offsets
    uint32_t* total_out;          // grand total (raw isect count)
    uint32_t seg_len;             // batched: ranks per view
    uint32_t view_shift;          // batched: key = view << view_shift | local tile
    const BatchView* views;
    uint32_t* ticket;             // zeroed before launch
    uint64_t* status;             // one word per workgroup, zeroed before launch
};

// Fixed-point reciprocal division for the bbox walk: q = li / wx for
// li < 2^22 from one float multiply and one correction (the float quotient is
// within one of the true one there); larger li takes the integer divide.
__device__ __forceinline__ uint32_t div_small(uint32_t li, uint32_t wx, float inv_wx) {
    if (li >= (1u << 22)) return li / wx;
    uint32_t q = (uint32_t)((float)li * inv_wx);
    const int r = (int)li - (int)(q * wx);
    if (r < 0) --q;
    else if (r >= (int)wx) ++q;
    return q;
}

template <bool ORDERED>
__global__ __launch_bounds__(kEmitThreads) void k_scan_emit(EmitArgs a) {
    __shared__ uint32_t s_incl[kEmitRanks];  // inclusive prefix within the wave's 256 ranks
    __shared__ uint2 s_bb[kEmitRanks];       // (min_x | min_y << 16, max_x | max_y << 16)
    __shared__ float s_inv[kEmitRanks];      // 1 / wx
    __shared__ uint32_t s_g[kEmitRanks];
    __shared__ uint32_t s_tmp[kEmitThreads / 64 + 1];
    __shared__ uint32_t s_bid, s_excl;

    const uint32_t t = threadIdx.x;
    if (t == 0) s_bid = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const uint32_t bid = s_bid;
    const uint32_t r0 = bid * kEmitRanks + t * kEmitRpt;
    const bool full = r0 + kEmitRpt <= a.n;

    // This thread's 4 consecutive ranks: splat and tile bbox of each.
    uint32_t gg[kEmitRpt], cnt[kEmitRpt];
    uint2 bbs4[kEmitRpt];
    if constexpr (ORDERED) {
        uint2 oc[kEmitRpt];
        if (full) {
            const uint4 p0 = reinterpret_cast<const uint4*>(a.ordered)[r0 / 2];
            const uint4 p1 = reinterpret_cast<const uint4*>(a.ordered)[r0 / 2 + 1];
            oc[0] = make_uint2(p0.x, p0.y); oc[1] = make_uint2(p0.z, p0.w);
            oc[2] = make_uint2(p1.x, p1.y); oc[3] = make_uint2(p1.z, p1.w);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kEmitRpt; ++k)
                oc[k] = r0 + k < a.n ? a.ordered[r0 + k] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (uint32_t k = 0; k < kEmitRpt; ++k) {
            gg[k] = oc[k].x;
            const uint32_t min_x = oc[k].y & 0xffu, min_y = (oc[k].y >> 8) & 0xffu,
                           max_x = (oc[k].y >> 16) & 0xffu, max_y = oc[k].y >> 24;
            cnt[k] = (max_x - min_x) * (max_y - min_y);
            bbs4[k] = make_uint2(min_x | (min_y << 16), max_x | (max_y << 16));
        }
        if (a.order_out) {
            if (full) {
                reinterpret_cast<uint4*>(a.order_out)[r0 / 4] = make_uint4(gg[0], gg[1], gg[2], gg[3]);
            } else {
#pragma unroll
                for (uint32_t k = 0; k < kEmitRpt; ++k)
                    if (r0 + k < a.n) a.order_out[r0 + k] = gg[k];
            }
        }
    } else {
        if (full) {
            const uint4 o = reinterpret_cast<const uint4*>(a.order)[r0 / 4];
            gg[0] = o.x; gg[1] = o.y; gg[2] = o.z; gg[3] = o.w;
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kEmitRpt; ++k) gg[k] = r0 + k < a.n ? a.order[r0 + k] : 0u;
        }
        size_t gi[kEmitRpt];
#pragma unroll
        for (uint32_t k = 0; k < kEmitRpt; ++k) {
            gi[k] = gg[k];
            if (a.seg_len) gi[k] = (size_t)((r0 + k) / a.seg_len) * a.seg_len + gg[k];
            cnt[k] = r0 + k < a.n ? a.tile_counts[gi[k]] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kEmitRpt; ++k)
            bbs4[k] = cnt[k] ? a.bbox[gi[k]] : make_uint2(0u, 0u);
    }
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kEmitRpt; ++k) {
        sum += cnt[k];
        const uint32_t wx = (bbs4[k].y & 0xffffu) - (bbs4[k].x & 0xffffu);
        s_bb[t * kEmitRpt + k] = bbs4[k];
        s_inv[t * kEmitRpt + k] = wx ? 1.0f / (float)wx : 0.0f;
        s_g[t * kEmitRpt + k] = gg[k];
    }

    // Workgroup exclusive prefix of the per-thread sums (u32, wrapping as
    // the reference's u32 scan does).
    uint32_t total;
    const uint32_t texcl = block_excl_scan_u32<kEmitThreads>(sum, s_tmp, &total);

    // Decoupled look-back for the prefix of everything before this workgroup.
    if (t == 0) {
        uint32_t excl = 0;
        if (bid == 0) {
            status_store(&a.status[0], kFlagPrefix, total);
        } else {
            status_store(&a.status[bid], kFlagAggregate, total);
            for (uint32_t p = bid - 1;;) {
                const uint64_t w = status_load(&a.status[p]);
                const uint32_t flag = (uint32_t)(w >> 32);
                if (flag == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)w;
                if (flag == kFlagPrefix) break;
                --p;  // an aggregate: keep walking (block 0 always holds a prefix)
            }
            status_store(&a.status[bid], kFlagPrefix, excl + total);
        }
        s_excl = excl;
        if (bid * kEmitRanks + kEmitRanks >= a.n && a.total_out) *a.total_out = excl + total;
    }
    __syncthreads();
    const uint32_t base = s_excl;

    // Per-rank offsets, and the inclusive prefix of each rank within its
    // wave's 256 ranks for the emission.
    const uint32_t lane = lane_id();
    const uint32_t wave_first = (t / 64) * 64;  // first thread of this wave
    const uint32_t wave_base_excl = __shfl(texcl, 0, 64);
    uint32_t run = texcl;
    uint32_t offs[kEmitRpt];
#pragma unroll
    for (uint32_t k = 0; k < kEmitRpt; ++k) {
        offs[k] = base + run;
        run += cnt[k];
        s_incl[t * kEmitRpt + k] = run - wave_base_excl;
    }
    if (a.offsets) {
        if (full) {
            reinterpret_cast<uint4*>(a.offsets)[r0 / 4] = make_uint4(offs[0], offs[1], offs[2], offs[3]);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kEmitRpt; ++k)
                if (r0 + k < a.n) a.offsets[r0 + k] = offs[k];
        }
    }
    // Wave range: [base + wave_base_excl, + wave_total).
    const uint32_t wave_total = __shfl(run, 63, 64) - wave_base_excl;
    const uint32_t out0 = base + wave_base_excl;
    const uint32_t* incl = s_incl + wave_first * kEmitRpt;
    const uint2* bbs = s_bb + wave_first * kEmitRpt;
    const float* invs = s_inv + wave_first * kEmitRpt;
    const uint32_t* gs = s_g + wave_first * kEmitRpt;
    const uint32_t rbase = bid * kEmitRanks + wave_first * kEmitRpt;
    __builtin_amdgcn_wave_barrier();
    // Four slots per lane per step (64 apart, so each store instruction still
    // covers 64 consecutive slots): four independent owner searches keep
    // four LDS round trips in flight.
    constexpr uint32_t Q = 4;
    for (uint32_t e0 = lane; e0 < wave_total; e0 += 64 * Q) {
        uint32_t lo[Q];
#pragma unroll
        for (uint32_t q = 0; q < Q; ++q) lo[q] = 0;
        // Owner: first k with incl[k] > e (256 entries, 8 halvings).
#pragma unroll
        for (uint32_t step = 128; step > 0; step >>= 1) {
#pragma unroll
            for (uint32_t q = 0; q < Q; ++q)
                if (incl[lo[q] + step - 1] <= e0 + 64 * q) lo[q] += step;
        }
#pragma unroll
        for (uint32_t q = 0; q < Q; ++q) {
            const uint32_t e = e0 + 64 * q;
            const uint32_t slot = out0 + e;
            if (e >= wave_total || slot >= a.max_isects) continue;
            const uint32_t k = lo[q];
            const uint32_t before = k ? incl[k - 1] : 0u;
            const uint32_t li = e - before;
            const uint2 bb = bbs[k];
            const uint32_t min_x = bb.x & 0xffffu, min_y = bb.x >> 16, max_x = bb.y & 0xffffu;
            const uint32_t wx = max_x - min_x;
            const uint32_t dy = div_small(li, wx, invs[k]), dx = li - dy * wx;
            if (a.seg_len) {
                // Batched: the key keeps the view above the local tile (only
                // the local tile bits are sorted; the input is already
                // view-major), the value is the view-major splat slot.
                const uint32_t v = (rbase + k) / a.seg_len;
                const uint32_t tpr = a.views[v].tiles_x;
                a.tile_ids[slot] = (v << a.view_shift) | ((min_x + dx) + (min_y + dy) * tpr);
                a.gaussian_ids[slot] = v * a.seg_len + gs[k];
            } else {
                a.tile_ids[slot] = (min_x + dx) + (min_y + dy) * a.tiles_per_row;
                a.gaussian_ids[slot] = gs[k];
            }
        }
    }
}

}  // namespace

uint32_t scan_emit_status_words(uint32_t n) { return 2u + 2u * div_up(n ? n : 1, kEmitRanks); }

hipError_t launch_scan_emit(hipStream_t s, const EmitSource& src, uint32_t n,
                            uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                            uint32_t* gaussian_ids, uint32_t* offsets, uint32_t* total_out,
                            uint32_t* status_words, uint32_t seg_len, const BatchView* views,
                            uint32_t view_shift) {
    if (n == 0) return hipSuccess;
    const uint32_t nb = div_up(n, kEmitRanks);
    hipError_t e = launch_fill_u32(s, status_words, 0u, scan_emit_status_words(n));
    if (e != hipSuccess) return e;
    EmitArgs a;
    a.order = src.order;
    a.tile_counts = src.tile_counts;
    a.bbox = reinterpret_cast<const uint2*>(src.bbox);
    a.ordered = src.ordered;
    a.order_out = src.order_out;
    a.n = n;
    a.tiles_per_row = tiles_per_row;
    a.max_isects = max_isects;
    a.tile_ids = tile_ids;
    a.gaussian_ids = gaussian_ids;
    a.offsets = offsets;
    a.total_out = total_out;
    a.seg_len = seg_len;
    a.view_shift = view_shift;
    a.views = views;
    a.ticket = status_words;
    a.status = reinterpret_cast<uint64_t*>(status_words + 2);
    if (src.ordered) k_scan_emit<true><<<nb, kEmitThreads, 0, s>>>(a);
    else k_scan_emit<false><<<nb, kEmitThreads, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace sfk
