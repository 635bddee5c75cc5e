// Binned pair emission straight into bucket order, for gfx950.
//
// The binned emission (sfk_bin.hip) writes one (supertile, splat | bbox)
// pair per depth rank and supertile its tile bbox meets, in rank order, then
// sorts the pairs by supertile with a radix pass.  A caller that rasterizes
// straight from the pairs (the batched EIG round) needs only every bucket's
// pairs -- bucket (local supertile s, view v), index s * views + v -- in rank
// order, and the bucket ranges.  This path writes them there directly, with
// no look-back, no pair sort and no ranges pass:
//
//  1. k_pair_count: per workgroup of 4096 ranks (all of one view), the isect
//     and pair totals and the pairs per local supertile;
//  2. k_pair_bases: one workgroup scans the totals into per-workgroup isect
//     and pair bases (the exclusive prefixes of the emission in rank order,
//     saturating at 2^31 - 1) and, when the pairs exceed the capacity, keeps
//     only those whose rank-order slot is below it, as the sorted path does
//     (such a pair's isects all lie at or past the capacity too);
//  3. an exclusive scan of the per-(supertile, workgroup) counts in
//     supertile-major order gives every bucket's start (k_pair_starts);
//  4. k_pair_write: per workgroup, the isect offsets and truncation flags of
//     k_bin_scan, and every kept pair at its bucket's base for the workgroup
//     + its stable rank among the workgroup's pairs of that supertile.
//
// The result equals the sorted path's (k_bin_scan + stable pair argsort +
// ranges) pair for pair.
#include <algorithm>

#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

constexpr uint32_t kThreads = 1024;
constexpr uint32_t kRpt = 4;
constexpr uint32_t kRanks = kThreads * kRpt;
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kSat = 0x7fffffffu;

struct PairArgs {
    const uint2* ordered;   // (in-view splat, compact bbox) per depth rank
    uint32_t seg_len;       // ranks per view
    uint32_t wpv;           // workgroups per view: div_up(seg_len, kRanks)
    uint32_t W;             // workgroups: views * wpv
    uint32_t S;             // local supertiles per view (the largest view's)
    uint32_t V;             // views
    const BinView* views;   // device, per view, or nullptr: view0
    BinView view0;
    uint32_t cap;
    uint32_t* counts;       // S * W + 1: kept pairs per (supertile, workgroup), scanned in place
    uint32_t* wg;           // 4 * W: isect total, pair total, isect base, pair base
    uint32_t* sstart;       // S * V + 1
    uint2* pairs;           // cap
    uint32_t* offs_by_splat;
    uint32_t* order_out;    // optional
    uint32_t* offsets;      // optional
    uint32_t* counters;     // [0] raw isect total, [1] kept pairs
    uint32_t* host_total;   // optional
};

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t s = a + b;  // both at most 2^31 - 1
    return s > kSat ? kSat : s;
}

// Workgroup w's view and rank range [r0, r1).
__device__ __forceinline__ void wg_range(const PairArgs& a, uint32_t w, uint32_t& v, uint32_t& r0,
                                         uint32_t& r1) {
    v = w / a.wpv;
    const uint32_t vb = v * a.seg_len;
    r0 = vb + (w - v * a.wpv) * kRanks;
    r1 = min(r0 + kRanks, vb + a.seg_len);
}

// A rank's pairs: the supertiles of its compact tile bbox, row by row.
// Pair j has local supertile d0 + (j / swx) * supers_x + j % swx.
struct Span {
    uint32_t cnt;  // isects (tiles of the bbox)
    uint32_t sc;   // pairs (supertiles of the bbox)
    uint32_t d0, swx;
};
__device__ __forceinline__ Span rank_span(uint32_t cb, uint32_t supers_x, uint32_t sh) {
    const uint32_t x0 = cb & 0xffu, y0 = (cb >> 8) & 0xffu, x1 = (cb >> 16) & 0xffu,
                   y1 = cb >> 24;
    Span s;
    s.cnt = (x1 - x0) * (y1 - y0);
    s.sc = 0;
    s.d0 = 0;
    s.swx = 1;
    if (s.cnt) {
        const uint32_t sx0 = x0 >> sh, sy0 = y0 >> sh;
        s.swx = ((x1 - 1) >> sh) - sx0 + 1;
        s.sc = s.swx * (((y1 - 1) >> sh) - sy0 + 1);
        s.d0 = sy0 * supers_x + sx0;
    }
    return s;
}
// j / swx for j < 2^11, swx in [1, 32], by a multiply with
// inv = ceil(2^16 / swx): j * (inv * swx - 2^16) < 2^16 keeps it exact.
__device__ __forceinline__ uint32_t span_inv(uint32_t swx) { return (65535u + swx) / swx; }
__device__ __forceinline__ uint32_t pair_digit(uint32_t d0, uint32_t swx, uint32_t inv,
                                               uint32_t supers_x, uint32_t j) {
    const uint32_t dy = (j * inv) >> 16;
    return d0 + dy * supers_x + (j - dy * swx);
}

// Adds a rank's first `k` pairs to hist[supertile] (LDS atomics; the order
// of the adds does not matter; no division: the supertile steps along the
// bbox rows).
__device__ __forceinline__ void hist_pairs(const Span& s, uint32_t k, uint32_t supers_x,
                                           uint32_t* hist) {
    uint32_t d = s.d0, col = 0;
    for (uint32_t j = 0; j < k; ++j) {
        atomicAdd(&hist[d], 1u);
        if (++col == s.swx) {
            col = 0;
            d += supers_x - s.swx + 1;
        } else {
            ++d;
        }
    }
}

// 1. Totals and per-supertile pair counts of each workgroup's ranks (thread
// t takes ranks t, t + 1024, ..: every load is one coalesced run).
__global__ __launch_bounds__(kThreads) void k_pair_count(PairArgs a) {
    __shared__ uint32_t s_hist[kPairMaxDigits];
    __shared__ uint32_t s_red[2][kWaves];
    const uint32_t t = threadIdx.x, lane = lane_id(), wv = t / 64, w = blockIdx.x;
    uint32_t v, r0, r1;
    wg_range(a, w, v, r0, r1);
    const BinView vw = a.views ? a.views[v] : a.view0;
    for (uint32_t i = t; i < a.S; i += kThreads) s_hist[i] = 0;
    uint32_t cb[kRpt];
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const uint32_t r = r0 + k * kThreads + t;
        cb[k] = r < r1 ? a.ordered[r].y : 0u;  // bbox 0: no isects
    }
    __syncthreads();
    uint32_t ci = 0, cp = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const Span s = rank_span(cb[k], vw.supers_x, vw.super_shift);
        ci += s.cnt;
        cp += s.sc;
        hist_pairs(s, s.sc, vw.supers_x, s_hist);
    }
    ci = wave_sum_u32(ci);
    cp = wave_sum_u32(cp);
    if (lane == 0) {
        s_red[0][wv] = ci;
        s_red[1][wv] = cp;
    }
    __syncthreads();
    for (uint32_t i = t; i < a.S; i += kThreads) a.counts[(size_t)i * a.W + w] = s_hist[i];
    if (t == 0) {
        uint32_t ti = 0, tp = 0;  // at most 4096 * 255^2 < 2^32
        for (uint32_t k = 0; k < kWaves; ++k) {
            ti += s_red[0][k];
            tp += s_red[1][k];
        }
        a.wg[4 * w] = ti;
        a.wg[4 * w + 1] = tp;
    }
}

// Inclusive prefix of a u64 across the wave.
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// 2. One workgroup: per-workgroup isect and pair bases; the kept pairs when
// they exceed the capacity.
__global__ __launch_bounds__(kThreads) void k_pair_bases(PairArgs a) {
    __shared__ uint64_t s_w[2][kWaves];
    __shared__ uint32_t s_hist[kPairMaxDigits];
    __shared__ uint32_t s_tmp[kWaves + 1];
    __shared__ uint32_t s_cut;
    const uint32_t t = threadIdx.x, lane = lane_id(), wv = t / 64;
    uint64_t carry_i = 0, carry_p = 0;
    for (uint32_t c0 = 0; c0 < a.W; c0 += kThreads) {
        const uint32_t w = c0 + t;
        const uint64_t ti = w < a.W ? a.wg[4 * w] : 0u, tp = w < a.W ? a.wg[4 * w + 1] : 0u;
        const uint64_t ii = wave_incl_scan_u64(ti), ip = wave_incl_scan_u64(tp);
        if (lane == 63) {
            s_w[0][wv] = ii;
            s_w[1][wv] = ip;
        }
        __syncthreads();
        uint64_t bi = carry_i, bp = carry_p, si = 0, sp = 0;
        for (uint32_t k = 0; k < kWaves; ++k) {
            if (k < wv) {
                bi += s_w[0][k];
                bp += s_w[1][k];
            }
            si += s_w[0][k];
            sp += s_w[1][k];
        }
        bi += ii - ti;
        bp += ip - tp;
        if (w < a.W) {
            a.wg[4 * w + 2] = (uint32_t)min<uint64_t>(bi, kSat);
            a.wg[4 * w + 3] = (uint32_t)min<uint64_t>(bp, kSat);
        }
        carry_i += si;
        carry_p += sp;
        __syncthreads();
    }
    if (t == 0) {
        const uint32_t ti = (uint32_t)min<uint64_t>(carry_i, kSat);
        a.counters[0] = ti;
        a.counters[1] = (uint32_t)min<uint64_t>(carry_p, a.cap);
        if (a.host_total) *a.host_total = ti;
        s_cut = a.W;
    }
    if (carry_p <= a.cap) return;  // every pair kept (uniform)
    // The pairs at rank-order slots >= cap are dropped: no pair of a
    // workgroup whose base is past the capacity, and the first cap - base of
    // the one that straddles it, recounted here.
    __syncthreads();
    for (uint32_t w = t; w < a.W; w += kThreads) {
        const uint32_t pb = a.wg[4 * w + 3], tp = a.wg[4 * w + 1];
        if (pb < a.cap && a.cap - pb < tp) s_cut = w;
        if (pb >= a.cap)
            for (uint32_t d = 0; d < a.S; ++d) a.counts[(size_t)d * a.W + w] = 0;
    }
    for (uint32_t i = t; i < a.S; i += kThreads) s_hist[i] = 0;
    __syncthreads();
    const uint32_t wc = s_cut;
    if (wc >= a.W) return;
    uint32_t v, r0, r1;
    wg_range(a, wc, v, r0, r1);
    const BinView vw = a.views ? a.views[v] : a.view0;
    Span s[kRpt];
    uint32_t cp = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const uint32_t r = r0 + t * kRpt + k;  // consecutive: rank order
        s[k] = rank_span(r < r1 ? a.ordered[r].y : 0u, vw.supers_x, vw.super_shift);
        cp += s[k].sc;
    }
    uint32_t total;
    uint32_t run = block_excl_scan_u32<kThreads>(cp, s_tmp, &total);
    const uint32_t pb = a.wg[4 * wc + 3];
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const uint32_t slot = pb + run;  // < 2^31: pb < cap <= 2^30
        const uint32_t kept = slot < a.cap ? min(s[k].sc, a.cap - slot) : 0u;
        hist_pairs(s[k], kept, vw.supers_x, s_hist);
        run += s[k].sc;
    }
    __syncthreads();
    for (uint32_t i = t; i < a.S; i += kThreads) a.counts[(size_t)i * a.W + wc] = s_hist[i];
}

// 3. Bucket (s, v) starts at the scanned count of (s, view v's first
// workgroup); the total closes the last bucket.
__global__ __launch_bounds__(256) void k_pair_starts(PairArgs a) {
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nb = a.S * a.V;
    if (b > nb) return;
    if (b == nb) {
        a.sstart[b] = a.counts[(size_t)a.S * a.W];
        return;
    }
    const uint32_t d = b / a.V, v = b - d * a.V;
    a.sstart[b] = a.counts[(size_t)d * a.W + v * a.wpv];
}

// 4. Offsets, truncation flags and the pairs at their bucket positions:
// per wave and supertile, the kept pairs counted (LDS atomics), offset by the
// workgroup's base for the supertile and by the waves before (their ranks
// come first); then each wave walks its kept pairs in rank order, 64 at a
// time, ranked among equal supertiles by match-any.  (Staging whole windows
// of the workgroup's pairs in LDS to store each supertile's run contiguously
// was measured slower: 377 vs 249 us for the 20-view 512^2 round at 8 px
// tiles.)
template <uint32_t DB>
__global__ __launch_bounds__(kThreads) void k_pair_write(PairArgs a) {
    constexpr uint32_t kOwnWin = 512;  // pair slots per owner window of a wave
    // LDS kept below 80 KB (two workgroups per CU): the per-wave supertile
    // positions sized for this instantiation's 2^DB supertiles, the spans in
    // 16 bits with the reciprocals of their widths in a table.
    constexpr uint32_t DS = 1u << DB;
    __shared__ uint32_t s_incl[kRanks];  // inclusive kept-pair prefix within the wave's ranks
    __shared__ uint16_t s_span[kRanks];  // d0 | swx << 8
    __shared__ uint2 s_val[kRanks];      // pair value: (splat | may-truncate << 31, bbox)
    __shared__ uint8_t s_own[kWaves][kOwnWin];
    __shared__ uint32_t s_wc[kWaves][DS];  // per wave and supertile: next position
    __shared__ uint32_t s_inv[33];         // span_inv(swx), swx in [1, 32]
    __shared__ uint32_t s_tmp[kWaves + 1];
    const uint32_t t = threadIdx.x, lane = lane_id(), wv = t / 64, w = blockIdx.x;
    uint32_t v, r0, r1;
    wg_range(a, w, v, r0, r1);
    const BinView vw = a.views ? a.views[v] : a.view0;
    const uint32_t rb = r0 + t * kRpt;
    uint2 oc[kRpt];
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k)
        oc[k] = rb + k < r1 ? a.ordered[rb + k] : make_uint2(0u, 0u);
    for (uint32_t i = t; i < kWaves * DS; i += kThreads) (&s_wc[0][0])[i] = 0;
    if (t >= 1 && t <= 32) s_inv[t] = span_inv(t);
    if (a.order_out) {
#pragma unroll
        for (uint32_t k = 0; k < kRpt; ++k)
            if (rb + k < r1) a.order_out[rb + k] = oc[k].x;
    }
    Span s[kRpt];
    uint32_t ci = 0, cp = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        s[k] = rank_span(oc[k].y, vw.supers_x, vw.super_shift);
        ci += s[k].cnt;
        cp += s[k].sc;
    }
    uint32_t tot;
    uint32_t irun = block_excl_scan_u32<kThreads>(ci, s_tmp, &tot);
    uint32_t prun = block_excl_scan_u32<kThreads>(cp, s_tmp, &tot);
    const uint32_t ib = a.wg[4 * w + 2], pb = a.wg[4 * w + 3];
    uint32_t kept[kRpt];
    uint32_t ck = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const uint32_t offs = sat_add(ib, irun);
        // A rank whose isects may reach the capacity carries a flag and
        // leaves its offset at its splat (raster.rs:12-48 drops slots >= cap).
        const bool cut = s[k].cnt && offs + s[k].cnt > a.cap;
        if (cut) a.offs_by_splat[vw.gid_add + oc[k].x] = offs;
        if (a.offsets && rb + k < r1) a.offsets[rb + k] = offs;
        const uint32_t slot = sat_add(pb, prun);
        kept[k] = slot < a.cap ? min(s[k].sc, a.cap - slot) : 0u;
        ck += kept[k];
        s_val[t * kRpt + k] = make_uint2(oc[k].x | (cut ? 0x80000000u : 0u), oc[k].y);
        s_span[t * kRpt + k] = (uint16_t)(s[k].d0 | (s[k].swx << 8));
        irun += s[k].cnt;
        prun += s[k].sc;
    }
    // Kept-pair prefix within the wave.
    const uint32_t incl = wave_incl_scan_u32(ck);
    {
        uint32_t run = incl - ck;
#pragma unroll
        for (uint32_t k = 0; k < kRpt; ++k) {
            run += kept[k];
            s_incl[t * kRpt + k] = run;
        }
    }
    __syncthreads();  // s_wc cleared
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) hist_pairs(s[k], kept[k], vw.supers_x, s_wc[wv]);
    __syncthreads();
    for (uint32_t d = t; d < a.S; d += kThreads) {
        uint32_t run = a.counts[(size_t)d * a.W + w];
        for (uint32_t k = 0; k < kWaves; ++k) {
            const uint32_t c = s_wc[k][d];
            s_wc[k][d] = run;
            run += c;
        }
    }
    __syncthreads();
    // The owning rank of each pair slot of a window by scatter (one LDS write
    // per pair).
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    const uint32_t first = wv * 64 * kRpt;
    const uint32_t* wincl = s_incl + first;
    uint32_t* wc = s_wc[wv];
    for (uint32_t w0 = 0; w0 < total; w0 += kOwnWin) {
        const uint32_t w1 = min(total, w0 + kOwnWin);
#pragma unroll
        for (uint32_t k = 0; k < kRpt; ++k) {
            const uint32_t hi = min(wincl[lane * kRpt + k], w1);
            for (uint32_t q = max(wincl[lane * kRpt + k] - kept[k], w0); q < hi; ++q)
                s_own[wv][q - w0] = (uint8_t)(lane * kRpt + k);
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e0 = w0; e0 < w1; e0 += 64) {
            const uint32_t e = e0 + lane;
            const bool live = e < w1;
            uint32_t d = 0, own = 0;
            if (live) {
                own = s_own[wv][e - w0];
                const uint32_t j = e - (own ? wincl[own - 1] : 0u);
                const uint32_t sp = s_span[first + own];
                const uint32_t swx = sp >> 8;
                d = pair_digit(sp & 0xffu, swx, s_inv[swx], vw.supers_x, j);
            }
            const uint64_t peers = match_peers<DB>(d, live);
            const uint32_t base = live ? wc[d] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (live) {
                const uint32_t rk = mask_rank(peers);
                a.pairs[base + rk] = s_val[first + own];
                if (rk == 0) wc[d] = base + (uint32_t)__popcll(peers);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

template <uint32_t DB>
hipError_t launch_pairs(hipStream_t st, const PairArgs& a, uint32_t* scan_sums) {
    k_pair_count<<<a.W, kThreads, 0, st>>>(a);
    k_pair_bases<<<1, kThreads, 0, st>>>(a);
    hipError_t e = launch_scan(st, nullptr, a.counts, a.counts, a.S * a.W, scan_sums, 0,
                               a.counts + (size_t)a.S * a.W);
    if (e != hipSuccess) return e;
    k_pair_starts<<<div_up(a.S * a.V + 1, 256), 256, 0, st>>>(a);
    k_pair_write<DB><<<a.W, kThreads, 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace

uint32_t pair_table_words(uint32_t n_per_view, uint32_t views, uint32_t digits) {
    const uint64_t W = (uint64_t)views * div_up(n_per_view, kRanks);
    return (uint32_t)(W * digits + 1 + 4 * W);
}

uint32_t pair_table_sum_words(uint32_t n_per_view, uint32_t views, uint32_t digits) {
    const uint64_t W = (uint64_t)views * div_up(n_per_view, kRanks);
    return div_up(W * digits, kScanBlock);
}

int bin_pairs_direct(sfk_ctx* c, const BinJob& j, uint32_t digits) {
    uint32_t* pk = nullptr;
    uint2* pv = nullptr;
    if (radix_scratch_spare_u32x2(j.sort, &pk, &pv))
        return fail(c, SFK_E_ARG, "binned pairs: pair scratch");
    const uint32_t seg = j.seg_len ? j.seg_len : j.n;
    PairArgs a;
    a.ordered = j.ordered;
    a.seg_len = seg;
    a.wpv = div_up(seg, kRanks);
    a.W = j.nviews * a.wpv;
    a.S = digits;
    a.V = j.nviews;
    a.views = j.views;
    a.view0 = j.view0;
    a.cap = j.cap;
    a.counts = j.pair_table;
    a.wg = j.pair_table + (size_t)digits * a.W + 1;
    a.sstart = j.sstart;
    a.pairs = pv;
    a.offs_by_splat = j.offs_by_splat;
    a.order_out = j.order_out;
    a.offsets = j.offsets;
    a.counters = j.counters;
    a.host_total = j.host_total;
    SFK_PROF(c, "bin_pairs");
    hipError_t e;
    if (digits <= 16) e = launch_pairs<4>(c->stream, a, j.pair_table_sums);
    else if (digits <= 64) e = launch_pairs<6>(c->stream, a, j.pair_table_sums);
    else e = launch_pairs<8>(c->stream, a, j.pair_table_sums);
    if (e != hipSuccess) return hip_fail(c, e, "binned pairs");
    *j.pairs_out = pv;
    return SFK_OK;
}

}  // namespace sfk
