// Binned isect emission for gfx950 on compact tile grids (at most 255 x 255
// tiles per view).  It yields exactly what the reference's isect stage
// yields -- the gather scan of tile counts in depth order (scan.rs:215-255),
// map_isects (raster.rs:12-48), which drops every isect whose emission slot
// reaches max_isects, the stable radix argsort by tile id and the tile
// ranges (raster.rs:52-65), as orchestrated by render.rs:271-364 -- namely,
// for every tile, the splats whose tile bbox covers it in depth-rank order
// and the tile's range in the isect arrays, but without sorting the isects:
//
//  1. k_bin_scan, one pass over the depth ranks: a decoupled look-back gives
//     two exclusive prefixes per rank, the isect offset of the reference's
//     emission (it decides the truncation) and the offset of the rank's
//     (supertile, splat | bbox) pairs, which it writes.  A supertile is a
//     block of 8 x 8 tiles; a bbox meets few of them (1.6 per splat for a
//     1M-splat 1080p frame, against 9.4 tiles).
//  2. a stable radix argsort of the pairs by supertile id (one pass up to 256
//     supertiles): every supertile's splats and bboxes in depth order, read
//     coalesced by the tile passes (no gather by rank).
//  3. each supertile's range of sorted pairs (k_tile_starts), cut into
//     chunks of kBinChunk pairs, one wavefront each (k_bin_chunk_map).
//  4. k_bin_chunks<false>: per chunk and per tile of its supertile, the isect
//     count.  A wavefront takes 64 pairs a round; lane i builds the 64-bit
//     mask of the supertile's tiles that splat i covers (minus the ones the
//     truncation drops), and a 64 x 64 bit-matrix transpose hands lane t the
//     mask of the round's splats that cover tile t.
//  5. k_bin_chunk_scan: per supertile and tile, the exclusive prefix over the
//     supertile's chunks and the tile's total; an exclusive scan of the
//     totals in tile order then gives tile_starts.
//  6. k_bin_chunks<true>: the same rounds again; lane t writes tile t's
//     isects of the round in rank order, at tile start + chunk prefix + its
//     running count.
//
// Every tile's list therefore holds the same splats in the same order as the
// sorted emission, and the arrays match it word for word.  The per-rank
// offsets saturate at 2^31 - 1 where the reference's u32 scan would wrap
// (both far beyond any isect capacity).
#include <algorithm>

#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

// Workgroups of 1024 threads, 4 ranks each: the ticket counter (one device
// atomic per workgroup on one word) and the look-back chain see 4096 ranks
// per workgroup; at 1024-rank workgroups the ticket alone cost ~10 us per
// million ranks on MI355X.
constexpr uint32_t kThreads = 1024;
constexpr uint32_t kRpt = 4;  // ranks per thread
constexpr uint32_t kRanks = kThreads * kRpt;
constexpr uint32_t kOwnWin = 512;  // pair slots per owner window of a wave (2 workgroups per CU)
constexpr uint32_t kSat = 0x7fffffffu;
constexpr uint64_t kFlagAggregate = 1, kFlagPrefix = 2;
constexpr uint32_t kRounds = kBinChunk / 64;

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t s = a + b;  // both at most 2^31 - 1: no u32 wrap
    return s > kSat ? kSat : s;
}

// Look-back status word: flag (2 bits) | pair prefix (31) | isect prefix (31).
__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t* p, uint64_t flag, uint32_t isects,
                                             uint32_t pairs) {
    __hip_atomic_store(p, (flag << 62) | ((uint64_t)pairs << 31) | isects, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ BinView view_of(const BinView* views, const BinView& view0,
                                           uint32_t v) {
    return views ? views[v] : view0;
}


struct BinScanArgs {
    const uint2* ordered;
    uint32_t n, seg_len;
    const BinView* views;
    BinView view0;
    uint32_t cap;
    uint32_t* order_out;
    uint32_t* offsets;
    uint32_t* offs_by_splat;
    uint32_t* pair_keys;
    uint2* pair_vals;
    uint32_t* counters;
    uint32_t* host_total;
    uint32_t* ticket;
    uint64_t* status;
};

// Workgroups of 1024 threads own 4096 consecutive ranks (4 per thread);
// logical ids come from a ticket, so every workgroup a look-back waits on has
// started and publishes its aggregate before it waits itself.  Each
// wavefront then writes the pairs of its 256 ranks (a contiguous range),
// consecutive lanes on consecutive slots, finding each slot's rank by binary
// search over the wave's inclusive pair prefix in LDS.
__global__ __launch_bounds__(kThreads) void k_bin_scan(BinScanArgs a) {
    __shared__ uint32_t s_incl[kRanks];  // inclusive pair prefix within the wave's ranks
    // Per rank: key of its first supertile (18 bits) | supertile columns
    // spanned << 18 | view supertiles per row << 24.
    __shared__ uint32_t s_span[kRanks];
    __shared__ uint2 s_val[kRanks];      // pair value: (splat | may-truncate << 31, bbox)
    // Per wave: the owning rank (index among the wave's 256) of each pair slot
    // of the current window of kOwnWin slots.
    __shared__ uint8_t s_own[kThreads / 64][kOwnWin];
    __shared__ uint32_t s_wi[kThreads / 64], s_wp[kThreads / 64];
    __shared__ uint32_t s_bid, s_ei, s_ep;

    const uint32_t t = threadIdx.x, lane = lane_id(), w = t / 64;
    if (t == 0) s_bid = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const uint32_t bid = s_bid;
    const uint32_t r0 = bid * kRanks + t * kRpt;
    const bool full = r0 + kRpt <= a.n;

    uint2 oc[kRpt];
    if (full) {
        const uint4 p0 = reinterpret_cast<const uint4*>(a.ordered)[r0 / 2];
        const uint4 p1 = reinterpret_cast<const uint4*>(a.ordered)[r0 / 2 + 1];
        oc[0] = make_uint2(p0.x, p0.y); oc[1] = make_uint2(p0.z, p0.w);
        oc[2] = make_uint2(p1.x, p1.y); oc[3] = make_uint2(p1.z, p1.w);
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kRpt; ++k)
            oc[k] = r0 + k < a.n ? a.ordered[r0 + k] : make_uint2(0u, 0u);
    }
    if (a.order_out) {
        if (full) {
            reinterpret_cast<uint4*>(a.order_out)[r0 / 4] =
                make_uint4(oc[0].x, oc[1].x, oc[2].x, oc[3].x);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kRpt; ++k)
                if (r0 + k < a.n) a.order_out[r0 + k] = oc[k].x;
        }
    }
    // View of the first rank (batched); later ranks step over boundaries.
    uint32_t v = a.seg_len ? (r0 < a.n ? r0 : a.n - 1) / a.seg_len : 0u;
    // The view's record is loaded once (pinned, so that it is not re-issued
    // inside the per-rank branches below) and again only past a boundary.
    BinView vw = view_of(a.views, a.view0, v);
    asm volatile("" : "+v"(vw.key_base), "+v"(vw.supers_x), "+v"(vw.gid_add));
    uint32_t cnt[kRpt], sc[kRpt], gadd[kRpt];
    uint32_t ci = 0, cp = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        const uint32_t r = r0 + k;
        if (a.seg_len && r < a.n && r >= (v + 1) * a.seg_len) {
            ++v;
            vw = view_of(a.views, a.view0, v);
        }
        const uint32_t cb = oc[k].y;
        const uint32_t x0 = cb & 0xffu, y0 = (cb >> 8) & 0xffu, x1 = (cb >> 16) & 0xffu,
                       y1 = cb >> 24;
        cnt[k] = (x1 - x0) * (y1 - y0);
        sc[k] = 0;
        uint32_t key0 = 0, span = 0;
        gadd[k] = 0;
        if (cnt[k]) {
            const uint32_t sx0 = x0 / kBinSuper, sy0 = y0 / kBinSuper;
            const uint32_t swx = (x1 - 1) / kBinSuper - sx0 + 1;
            const uint32_t swy = (y1 - 1) / kBinSuper - sy0 + 1;
            sc[k] = swx * swy;
            key0 = vw.key_base + sy0 * vw.supers_x + sx0;
            span = swx | (vw.supers_x << 6);
            gadd[k] = vw.gid_add;
        }
        s_span[t * kRpt + k] = key0 | (span << 18);
        ci += cnt[k];
        cp += sc[k];
    }

    // Workgroup exclusive prefixes of both per-thread sums.
    const uint32_t inc_i = wave_incl_scan_u32(ci), inc_p = wave_incl_scan_u32(cp);
    if (lane == 63) {
        s_wi[w] = inc_i;
        s_wp[w] = inc_p;
    }
    __syncthreads();
    uint32_t wbase_i = 0, wbase_p = 0, tot_i = 0, tot_p = 0;
#pragma unroll
    for (uint32_t k = 0; k < kThreads / 64; ++k) {
        if (k < w) {
            wbase_i += s_wi[k];
            wbase_p += s_wp[k];
        }
        tot_i += s_wi[k];
        tot_p += s_wp[k];
    }
    const uint32_t texcl_i = wbase_i + inc_i - ci, texcl_p = wbase_p + inc_p - cp;

    // Decoupled look-back over the workgroups before this one, by the first
    // wavefront: each step reads 64 predecessors' status words at once
    // (lane l: workgroup top - l), adds the values up to and including the
    // nearest prefix, and steps 64 further back only if none of them is one.
    if (w == 0) {
        uint32_t ei = 0, ep = 0;
        if (bid == 0) {
            if (lane == 0) status_store(&a.status[0], kFlagPrefix, tot_i, tot_p);
        } else {
            if (lane == 0) status_store(&a.status[bid], kFlagAggregate, tot_i, tot_p);
            for (int32_t top = (int32_t)bid - 1;;) {
                const int32_t p = top - (int32_t)lane;
                // Below workgroup 0: a zero prefix (never reached past
                // workgroup 0's own prefix, which is always published).
                const uint64_t sw = p >= 0 ? status_load(&a.status[p]) : (kFlagPrefix << 62);
                const uint32_t flag = (uint32_t)(sw >> 62);
                const uint64_t pre = __ballot(flag == kFlagPrefix);
                const uint64_t need = pre ? (2ull << __builtin_ctzll(pre)) - 1ull : ~0ull;
                if (__ballot(flag == 0) & need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;  // a needed predecessor has not published yet
                }
                const bool take = (need >> lane) & 1ull;
                uint64_t vi = take ? (sw & kSat) : 0ull, vp = take ? ((sw >> 31) & kSat) : 0ull;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    vi += __shfl_xor(vi, o, 64);
                    vp += __shfl_xor(vp, o, 64);
                }
                ei = (uint32_t)min<uint64_t>((uint64_t)ei + vi, kSat);
                ep = (uint32_t)min<uint64_t>((uint64_t)ep + vp, kSat);
                if (pre) break;
                top -= 64;
            }
            if (lane == 0)
                status_store(&a.status[bid], kFlagPrefix, sat_add(ei, tot_i), sat_add(ep, tot_p));
        }
        if (lane == 0) {
            s_ei = ei;
            s_ep = ep;
            if (bid * kRanks + kRanks >= a.n) {
                a.counters[0] = sat_add(ei, tot_i);
                a.counters[1] = min(sat_add(ep, tot_p), a.cap);
                if (a.host_total) *a.host_total = sat_add(ei, tot_i);
            }
        }
    }
    __syncthreads();
    const uint32_t ei = s_ei, ep = s_ep;

    // Per-rank isect offsets; the wave-local inclusive pair prefix.
    const uint32_t wave_excl_p = __shfl(texcl_p, 0, 64);
    uint32_t run_i = texcl_i, run_p = texcl_p;
    uint32_t offs[kRpt], phi[kRpt];
#pragma unroll
    for (uint32_t k = 0; k < kRpt; ++k) {
        offs[k] = sat_add(ei, run_i);
        run_i += cnt[k];
        run_p += sc[k];
        phi[k] = run_p - wave_excl_p;
        s_incl[t * kRpt + k] = phi[k];
        // A rank whose isects may reach the capacity carries a flag in its
        // pairs and leaves its offset at its splat: the only ranks whose
        // offset the tile passes need (raster.rs:12-48 drops slots >= cap).
        const bool cut = cnt[k] && offs[k] + cnt[k] > a.cap;
        s_val[t * kRpt + k] = make_uint2(oc[k].x | (cut ? 0x80000000u : 0u), oc[k].y);
        if (cut) a.offs_by_splat[gadd[k] + oc[k].x] = offs[k];
    }
    if (a.offsets) {
        if (full) {
            reinterpret_cast<uint4*>(a.offsets)[r0 / 4] =
                make_uint4(offs[0], offs[1], offs[2], offs[3]);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kRpt; ++k)
                if (r0 + k < a.n) a.offsets[r0 + k] = offs[k];
        }
    }

    // The wave's pairs: slots [pbase, pbase + wave_total) below the capacity.
    // A slot at or past the capacity holds a pair whose isects all lie at or
    // past it too (the j-th supertile of a bbox starts at tile index >= j),
    // so dropping it drops only truncated isects.
    const uint32_t wave_total = __shfl(run_p, 63, 64) - wave_excl_p;
    const uint32_t pbase = sat_add(ep, wave_excl_p);
    const uint32_t first = w * 64 * kRpt;
    const uint32_t* incl = s_incl + first;
    __builtin_amdgcn_wave_barrier();
    if (pbase >= a.cap) return;
    // Owners by scatter, a window of kOwnWin slots at a time: each rank writes
    // its index into the slots of its pairs, then each slot reads its owner
    // (the wave's ranks own consecutive slot ranges, so one write per pair
    // and one read replace a binary search over the wave's pair prefix).
    for (uint32_t w0 = 0; w0 < wave_total; w0 += kOwnWin) {
        if (pbase + w0 >= a.cap) break;
        const uint32_t w1 = min(wave_total, w0 + kOwnWin);
#pragma unroll
        for (uint32_t k = 0; k < kRpt; ++k) {
            const uint32_t hi = min(phi[k], w1);
            for (uint32_t q = max(phi[k] - sc[k], w0); q < hi; ++q)
                s_own[w][q - w0] = (uint8_t)(lane * kRpt + k);
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e = w0 + lane; e < w1; e += 64) {
            const uint32_t slot = pbase + e;
            if (slot >= a.cap) break;  // slots only grow with e
            const uint32_t k = s_own[w][e - w0];
            const uint32_t j = e - (k ? incl[k - 1] : 0u);
            const uint32_t span = s_span[first + k];
            const uint32_t key0 = span & 0x3ffffu;
            const uint32_t swx = (span >> 18) & 63u, row = span >> 24;
            const uint32_t dy = j < swx ? 0u : j / swx;
            a.pair_keys[slot] = key0 + dy * row + (j - dy * swx);
            a.pair_vals[slot] = s_val[first + k];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Chunk slots of each supertile: slot start / kBinChunk + s + k describes
// chunk k of supertile s as (view << 24 | s, first pair, end pair); the slots
// past the supertile's last chunk, up to the next supertile's first, are gaps
// (first >= end).
__global__ __launch_bounds__(256) void k_bin_chunk_map(const uint32_t* __restrict__ sstart,
                                                       uint32_t supers,
                                                       const BinView* __restrict__ views,
                                                       uint32_t nviews,
                                                       uint4* __restrict__ map) {
    // One workgroup per supertile, one thread per chunk: contiguous records.
    const uint32_t s = blockIdx.x;
    if (s >= supers) return;
    const uint32_t v = s % nviews;
    const uint32_t st = sstart[s], en = sstart[s + 1];
    const uint32_t c0 = st / kBinChunk + s, c1 = en / kBinChunk + s + 1;
    for (uint32_t c = c0 + threadIdx.x; c < c1; c += 256u) {
        const uint32_t p0 = st + (c - c0) * kBinChunk;
        map[c] = make_uint4((v << 24) | s, p0, min(p0 + kBinChunk, en), 0u);
    }
}

struct ChunkArgs {
    const uint2* pvals;     // sorted pair values: (splat | may-truncate << 31, bbox)
    const uint32_t* sstart;
    const uint4* map;
    const uint32_t* offs_by_splat;
    const BinView* views;
    BinView view0;
    uint32_t supers, cap, nviews;
    uint32_t* counts;              // count pass: per chunk per tile; place: chunk prefixes
    const uint32_t* tile_starts;   // place pass
    uint32_t* gids;
    uint32_t* tile_ids;
    uint32_t* tile_order;          // place pass, optional: tiles heaviest first
    uint32_t order_tiles;
    // Count pass, one view of at most 256 supertiles: the pair sort's digit
    // totals, from which every workgroup derives the supertile ranges and
    // chunk slots itself; workgroup 0 writes them out (sstart, map) for the
    // later passes.  Null: sstart and map are read.
    const uint32_t* totals;
    uint32_t* sstart_out;
    uint4* map_out;
    // Count pass, optional: two word ranges it zeroes (the emission scan's
    // look-back words, done with by then, and other counters), so that the
    // next emission needs no clearing launch.
    uint32_t* zero_a;
    uint32_t zero_a_n;
    uint32_t* zero_b;
    uint32_t zero_b_n;
};

// Supertile starts (exclusive prefixes of the pair counts, total at
// [supers]) into LDS by one 256-thread workgroup, supers <= 256; with `out`,
// also into device memory with the chunk map (the records of k_bin_chunk_map).
__device__ void supertile_ranges(const uint32_t* __restrict__ totals, uint32_t supers,
                                 uint32_t* s_start, uint32_t* tmp, uint32_t* __restrict__ sstart,
                                 uint4* __restrict__ map, bool out) {
    const uint32_t t = threadIdx.x;
    const uint32_t c = t < supers ? totals[t] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan_u32<256>(c, tmp, &total);
    if (t < supers) s_start[t] = ex;
    if (t == 0) s_start[supers] = total;
    __syncthreads();
    if (out) {
        if (t < supers) sstart[t] = ex;
        if (t == 0) sstart[supers] = total;
        if (t < supers) {
            const uint32_t st = ex, en = s_start[t + 1];
            const uint32_t c0 = st / kBinChunk + t, c1 = en / kBinChunk + t + 1;
            for (uint32_t k = c0; k < c1; ++k) {
                const uint32_t p0 = st + (k - c0) * kBinChunk;
                map[k] = make_uint4(t, p0, min(p0 + kBinChunk, en), 0u);
            }
        }
    }
}

// Chunk slot c from the supertile starts in LDS: (supertile, first pair,
// end pair), first >= end for a gap slot.  Slot c0(s) = start(s) / kBinChunk
// + s is supertile s's first; the last s with c0(s) <= c owns slot c.
__device__ __forceinline__ uint4 chunk_slot(const uint32_t* s_start, uint32_t supers, uint32_t c) {
    uint32_t lo = 0, hi = supers;  // c0(lo) <= c < c0(hi), c0(supers) = infinity
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_start[mid] / kBinChunk + mid <= c) lo = mid;
        else hi = mid;
    }
    const uint32_t st = s_start[lo], en = s_start[lo + 1];
    const uint32_t p0 = st + (c - (st / kBinChunk + lo)) * kBinChunk;
    return make_uint4(lo, p0, min(p0 + kBinChunk, en), 0u);
}

// Bits (ty - oy) * 8 + (tx - ox) of the supertile at tile (ox, oy) that the
// compact bbox covers and the truncation keeps.  Isect (tx, ty) of a rank
// with isect offset `off` has the emission slot
// off + (ty - y0) * (x1 - x0) + (tx - x0) (raster.rs:12-48 walks the bbox row
// by row); it is kept when that slot is below the capacity.
__device__ __forceinline__ uint64_t cover_mask(uint32_t cb, uint32_t ox, uint32_t oy, bool trunc,
                                               uint32_t off, uint32_t cap) {
    const int x0 = (int)(cb & 0xffu), y0 = (int)((cb >> 8) & 0xffu);
    const int x1 = (int)((cb >> 16) & 0xffu), y1 = (int)(cb >> 24);
    const int rx0 = max(x0 - (int)ox, 0), rx1 = min(x1 - (int)ox, 8);
    const int ry0 = max(y0 - (int)oy, 0), ry1 = min(y1 - (int)oy, 8);
    if (rx0 >= rx1 || ry0 >= ry1) return 0;
    const uint64_t row = (uint64_t)(((1u << (rx1 - rx0)) - 1u) << rx0);
    const uint64_t rep = 0x0101010101010101ull >> (8 * (8 - (ry1 - ry0)));
    uint64_t m = (row * rep) << (8 * ry0);
    if (trunc) {
        if (off >= cap) return 0;
        const uint32_t wx = (uint32_t)(x1 - x0), lim = cap - off;  // kept: index < lim
        if (lim < wx * (uint32_t)(y1 - y0)) {
            // Rows y0 .. y0 + qr - 1 whole, then qc tiles of row y0 + qr.
            const uint32_t qr = lim / wx, qc = lim - qr * wx;
            const int lr = y0 + (int)qr - (int)oy;
            uint64_t keep = lr <= 0 ? 0ull : lr >= 8 ? ~0ull : (1ull << (8 * lr)) - 1ull;
            if (lr >= 0 && lr < 8) {
                const int lc = min(max(x0 + (int)qc - (int)ox, 0), 8);
                keep |= (uint64_t)((1u << lc) - 1u) << (8 * lr);
            }
            m &= keep;
        }
    }
    return m;
}

// One wavefront per chunk of kBinChunk sorted pairs, looping over the chunk
// slots.  Lane t owns tile (ox + t % 8, oy + t / 8) of the chunk's supertile.
// The place pass stages the chunk's isects in LDS tile by tile, each as a
// 16-bit word (tile << 9 | round << 6 | lane of the splat), in groups of
// consecutive rounds holding at most kStage isects (a round holds at most
// 64 x 64), then copies every group out with consecutive lanes on
// consecutive staged words: a tile's isects of a group form one contiguous
// run of its range, so each store covers a few runs instead of 64 tiles.
constexpr uint32_t kStage = 4096;
static_assert(kRounds == 8 || kRounds == 4, "round totals use a transposed wave sum");
// Lane l of round_totals' result holds the wave total of round l / (64 / kRounds).
template <uint32_t N>
__device__ __forceinline__ uint32_t round_totals(const uint32_t (&cnt)[N]) {
    if constexpr (N == 8) return wave_sum8_transposed(cnt);
    else return wave_sum4_transposed(cnt);
}
constexpr uint32_t kRoundLane = 64 / kRounds;

// Work class of a tile with c isects: 0 for none, else 1 + 2 log2(c) + the
// bit below the leading one (classes a factor of ~1.4 apart), at most 64.
__device__ __forceinline__ uint32_t work_class(uint32_t c) {
    if (!c) return 0u;
    const uint32_t p = 31u - (uint32_t)__clz(c);
    const uint32_t sub = p ? (c >> (p - 1u)) & 1u : 0u;
    return 1u + 2u * p + sub;
}

// Lanes of the wave holding the same work class as this lane (7 ballots),
// restricted to valid lanes.
__device__ __forceinline__ uint64_t class_peers(uint32_t k, bool valid) {
    return match_peers<7>(k, valid);
}

// The rasterizers' launch order: the n tiles of `starts` (exclusive
// prefixes, total at [n]) heaviest work class first, by a counting sort on
// work_class of their isect counts (order inside a class unspecified).  One
// 256-thread workgroup; hist holds 65 words of LDS.
template <uint32_t NT = 256>
__device__ void tile_order_block(const uint32_t* __restrict__ starts, uint32_t n,
                                 uint32_t* __restrict__ order, uint32_t* hist) {
    constexpr uint32_t U = NT >= 1024 ? 2 : 8;  // tiles per thread whose counts load together
    for (uint32_t i = threadIdx.x; i < 65; i += NT) hist[i] = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < n; t0 += NT * U) {
        uint32_t k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t t = t0 + u * NT + threadIdx.x;
            k[u] = t < n ? starts[t + 1] - starts[t] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t t = t0 + u * NT + threadIdx.x;
            const bool valid = t < n;
            const uint32_t kc = work_class(k[u]);
            const uint64_t peers = class_peers(kc, valid);
            if (valid && mask_rank(peers) == 0) atomicAdd(&hist[kc], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 64; k >= 0; --k) {
            const uint32_t c = hist[k];
            hist[k] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t t0 = 0; t0 < n; t0 += NT * U) {
        uint32_t k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t t = t0 + u * NT + threadIdx.x;
            k[u] = t < n ? starts[t + 1] - starts[t] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t t = t0 + u * NT + threadIdx.x;
            const bool valid = t < n;
            const uint32_t kc = work_class(k[u]);
            const uint64_t peers = class_peers(kc, valid);
            const uint32_t rank = mask_rank(peers);
            uint32_t base = 0;
            if (valid && rank == 0) base = atomicAdd(&hist[kc], (uint32_t)__popcll(peers));
            const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : 0;
            base = __shfl(base, leader, 64);
            if (valid) order[base + rank] = t;
        }
    }
}

// With a.tile_order (place pass), workgroup 0 writes the tile order, off
// the critical path, while the others place isects.
template <bool PLACE>
__global__ __launch_bounds__(256) void k_bin_chunks(ChunkArgs a) {
    constexpr uint32_t SW = PLACE ? 4 : 1;
    __shared__ uint16_t s_st[SW][PLACE ? kStage : 1];
    __shared__ uint32_t s_gid[SW][PLACE ? kBinChunk : 1];
    __shared__ uint32_t s_start[PLACE ? 1 : 257];
    __shared__ uint32_t s_tmp[PLACE ? 1 : 5];
    uint32_t blk = blockIdx.x, nblk = gridDim.x;
    const bool own_ranges = !PLACE && a.totals;
    if constexpr (!PLACE) {
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < a.zero_a_n + a.zero_b_n;
             i += gridDim.x * 256u) {
            if (i < a.zero_a_n) a.zero_a[i] = 0u;
            else a.zero_b[i - a.zero_a_n] = 0u;
        }
        if (own_ranges)
            supertile_ranges(a.totals, a.supers, s_start, s_tmp, a.sstart_out, a.map_out,
                             blockIdx.x == 0);
    }
    if constexpr (PLACE) {
        if (a.tile_order) {
            if (blockIdx.x == 0) {
                tile_order_block(a.tile_starts, a.order_tiles, a.tile_order, &s_gid[0][0]);
                return;
            }
            blk -= 1;
            nblk -= 1;
        }
    }
    const uint32_t lane = lane_id(), w = threadIdx.x / 64u;
    const uint32_t wave = blk * 4u + w, nwaves = nblk * 4u;
    const uint32_t slots = (own_ranges ? s_start[a.supers] : a.sstart[a.supers]) / kBinChunk + a.supers;
    const TransposeLanes tl = transpose_lanes();
    for (uint32_t c = wave; c < slots; c += nwaves) {
        const uint4 d = own_ranges ? chunk_slot(s_start, a.supers, c) : a.map[c];
        const uint32_t p0 = d.y, pe = d.z;
        if (p0 >= pe) continue;  // a gap slot
        const uint32_t s = d.x & 0xffffffu;
        const uint32_t nr = (pe - p0 + 63) / 64;  // rounds in this chunk
        const BinView vw = view_of(a.views, a.view0, d.x >> 24);
        const uint32_t ls = s / a.nviews;
        const uint32_t sy = ls / vw.supers_x, sx = ls - sy * vw.supers_x;
        const uint32_t ox = sx * kBinSuper, oy = sy * kBinSuper;
        const uint32_t tx = ox + (lane & 7u), ty = oy + (lane >> 3);
        uint32_t base = 0;
        if constexpr (PLACE) {
            base = a.counts[(size_t)c * 64 + lane];
            if (tx < vw.tiles_x && ty < vw.tiles_y)
                base += a.tile_starts[(ty * vw.tiles_x + tx) * vw.out_mul + vw.out_add];
        }
        // All rounds' pairs first (coalesced), then the offsets of flagged ones.
        uint2 oc[kRounds];
        uint32_t off[kRounds];
#pragma unroll
        for (uint32_t k = 0; k < kRounds; ++k) {
            const uint32_t i = p0 + k * 64 + lane;
            oc[k] = i < pe ? a.pvals[i] : make_uint2(0u, 0u);  // bbox 0: no tiles
        }
#pragma unroll
        for (uint32_t k = 0; k < kRounds; ++k)
            off[k] = (oc[k].x >> 31) ? a.offs_by_splat[vw.gid_add + (oc[k].x & 0x7fffffffu)] : 0u;
        // Per round: lane t's mask of the round's splats covering tile t.
        uint64_t col[kRounds];
        uint32_t cnt[kRounds];
#pragma unroll
        for (uint32_t k = 0; k < kRounds; ++k) {
            col[k] = 0;
            if (k < nr) {
                const uint64_t mask =
                    cover_mask(oc[k].y, ox, oy, (oc[k].x >> 31) != 0, off[k], a.cap);
                col[k] = wave_transpose64(mask, tl);
            }
            cnt[k] = (uint32_t)__popcll(col[k]);
        }
        if constexpr (!PLACE) {
            uint32_t total = 0;
#pragma unroll
            for (uint32_t k = 0; k < kRounds; ++k) total += cnt[k];
            a.counts[(size_t)c * 64 + lane] = total;
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kRounds; ++k)
                s_gid[w][k * 64 + lane] = (oc[k].x & 0x7fffffffu) + vw.gid_add;
            const uint32_t tile = ty * vw.tiles_x + tx;
            // Round totals, then groups of consecutive rounds of at most
            // kStage isects (all wave-uniform).
            const uint32_t tk = round_totals(cnt);
            uint32_t grp[kRounds];
            uint32_t ng = 0, acc = 0;
#pragma unroll
            for (uint32_t k = 0; k < kRounds; ++k) {
                const uint32_t t = __builtin_amdgcn_readlane(tk, kRoundLane * k);
                if (k < nr) {
                    if (k > 0 && acc + t > kStage) {
                        ++ng;
                        acc = 0;
                    }
                    acc += t;
                }
                grp[k] = ng;
            }
            for (uint32_t q = 0; q <= ng; ++q) {
                uint32_t mine = 0;
#pragma unroll
                for (uint32_t k = 0; k < kRounds; ++k)
                    if (k < nr && grp[k] == q) mine += cnt[k];
                const uint32_t lo = wave_incl_scan_u32(mine) - mine;
                uint32_t gsum = 0;
#pragma unroll
                for (uint32_t k = 0; k < kRounds; ++k)
                    if (k < nr && grp[k] == q) gsum += __builtin_amdgcn_readlane(tk, kRoundLane * k);
                // Stage: tile t's isects of the group at lo.., in rank order.
                uint32_t pos = lo;
#pragma unroll
                for (uint32_t k = 0; k < kRounds; ++k) {
                    if (!(k < nr && grp[k] == q)) continue;
                    // The two 32-bit halves in turn: 32-bit bit scans.
#pragma unroll
                    for (uint32_t h = 0; h < 2; ++h) {
                        uint32_t cm = (uint32_t)(col[k] >> (32 * h));
                        const uint32_t tag = (lane << 9) | (k * 64 + 32 * h);
                        while (cm) {
                            s_st[w][pos++] = (uint16_t)(tag + __builtin_ctz(cm));
                            cm &= cm - 1;
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                // Copy out: staged word e of tile j goes to slot e + base_j - lo_j.
                const uint32_t delta = base - lo;
                for (uint32_t e0 = 0; e0 < gsum; e0 += 64) {
                    const uint32_t e = e0 + lane;
                    const uint32_t v = e < gsum ? (uint32_t)s_st[w][e] : 0u;
                    const int j = (int)(v >> 9);
                    const uint32_t slot = e + (uint32_t)__shfl((int)delta, j, 64);
                    const uint32_t tid = a.tile_ids ? (uint32_t)__shfl((int)tile, j, 64) : 0u;
                    if (e < gsum && slot < a.cap) {  // holds by construction; a guard
                        a.gids[slot] = s_gid[w][v & 511u];  // chunk index (9 bits)
                        if (a.tile_ids) a.tile_ids[slot] = tid;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                base += mine;
            }
        }
    }
}

// Exclusive scan of buf[0, n) in place with the total in buf[n] by one
// workgroup of NT threads, 16 consecutive words per thread per sweep.
constexpr uint32_t kSmallScanEpt = 16;
constexpr uint32_t kSmallScanMax = 1u << 16;
template <uint32_t NT>
__device__ void block_scan_small(uint32_t* __restrict__ buf, uint32_t n) {
    __shared__ uint32_t tmp[NT / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += NT * kSmallScanEpt) {
        const uint32_t first = c0 + threadIdx.x * kSmallScanEpt;
        const bool full = first + kSmallScanEpt <= n;
        uint32_t v[kSmallScanEpt];
        if (full) {
#pragma unroll
            for (uint32_t q = 0; q < kSmallScanEpt / 4; ++q) {
                const uint4 x = reinterpret_cast<const uint4*>(buf + first)[q];
                v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (uint32_t e = 0; e < kSmallScanEpt; ++e) v[e] = first + e < n ? buf[first + e] : 0u;
        }
        uint32_t s = 0;
#pragma unroll
        for (uint32_t e = 0; e < kSmallScanEpt; ++e) s += v[e];
        uint32_t total;
        uint32_t run = carry + block_excl_scan_u32<NT>(s, tmp, &total);
        uint32_t o[kSmallScanEpt];
#pragma unroll
        for (uint32_t e = 0; e < kSmallScanEpt; ++e) {
            o[e] = run;
            run += v[e];
        }
        if (full) {
#pragma unroll
            for (uint32_t q = 0; q < kSmallScanEpt / 4; ++q)
                reinterpret_cast<uint4*>(buf + first)[q] =
                    make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
        } else {
#pragma unroll
            for (uint32_t e = 0; e < kSmallScanEpt; ++e)
                if (first + e < n) buf[first + e] = o[e];
        }
        carry += total;
    }
    if (threadIdx.x == 0) buf[n] = carry;
}

// One workgroup per supertile: per tile (lane), the exclusive prefix of the
// chunk counts over the supertile's chunks (in place), and the tile total
// into tile_starts.  The four waves split the chunks; their partial sums meet
// in LDS.
// With class_count (65 words, zero on entry) and tile_cls (per tile), it
// also counts its tiles' work classes, one device atomic per distinct class
// of the supertile, and gives each tile its class and index within the
// class (class << 24 | index; the order inside a class is the atomics').
__global__ __launch_bounds__(256) void k_bin_chunk_scan(const uint32_t* __restrict__ sstart,
                                                        const BinView* __restrict__ views,
                                                        BinView view0, uint32_t nviews,
                                                        uint32_t* __restrict__ counts,
                                                        uint32_t* __restrict__ tile_starts,
                                                        uint32_t* __restrict__ class_count,
                                                        uint32_t* __restrict__ tile_cls) {
    __shared__ uint32_t s_part[4][64];
    const uint32_t s = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u, g = threadIdx.x >> 6;
    const uint32_t st = sstart[s], en = sstart[s + 1];
    const uint32_t c0 = st / kBinChunk + s;
    const uint32_t nch = (en - st + kBinChunk - 1) / kBinChunk;
    const uint32_t q = (nch + 3) / 4;
    const uint32_t k0 = min(g * q, nch), k1 = min(k0 + q, nch);
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t k = k0; k < k1; ++k) sum += counts[(size_t)(c0 + k) * 64 + lane];
    s_part[g][lane] = sum;
    __syncthreads();
    uint32_t run = 0, tot = 0;
#pragma unroll
    for (uint32_t h = 0; h < 4; ++h) {
        const uint32_t p = s_part[h][lane];
        if (h < g) run += p;
        tot += p;
    }
#pragma unroll 8
    for (uint32_t k = k0; k < k1; ++k) {
        const size_t i = (size_t)(c0 + k) * 64 + lane;
        const uint32_t cnt = counts[i];
        counts[i] = run;
        run += cnt;
    }
    if (g == 0) {
        const BinView vw = view_of(views, view0, s % nviews);
        const uint32_t ls = s / nviews;
        const uint32_t sy = ls / vw.supers_x, sx = ls - sy * vw.supers_x;
        const uint32_t tx = sx * kBinSuper + (lane & 7u), ty = sy * kBinSuper + (lane >> 3);
        const bool live = tx < vw.tiles_x && ty < vw.tiles_y;
        const uint32_t slot = (ty * vw.tiles_x + tx) * vw.out_mul + vw.out_add;
        if (live) tile_starts[slot] = tot;
        if (class_count) {
            const uint32_t kc = work_class(tot);
            const uint64_t peers = class_peers(kc, live);
            const uint32_t rank = mask_rank(peers);
            uint32_t base = 0;
            if (live && rank == 0) base = atomicAdd(&class_count[kc], (uint32_t)__popcll(peers));
            base = __shfl(base, live ? __ffsll((unsigned long long)peers) - 1 : 0, 64);
            if (live) tile_cls[slot] = (kc << 24) | (base + rank);
        }
    }
}

// Exclusive scan of buf[0, n) in place with the total in buf[n]: one
// workgroup.
constexpr uint32_t kSmallScanThreads = 1024;
__global__ __launch_bounds__(kSmallScanThreads) void k_scan_small(uint32_t* __restrict__ buf,
                                                                   uint32_t n) {
    block_scan_small<kSmallScanThreads>(buf, n);
}

// The same scan of the tile counts into tile_starts, then the rasterizers'
// heaviest-first tile order (when no placement pass writes it) from the
// classes the chunk scan counted: class starts heaviest first (one wave),
// then every tile to its class start + its index within the class.  The
// class counts are left zero for the next prepare.
__global__ __launch_bounds__(kSmallScanThreads) void k_scan_small_order(
    uint32_t* __restrict__ buf, uint32_t n, uint32_t* __restrict__ order,
    uint32_t* __restrict__ class_count, const uint32_t* __restrict__ tile_cls) {
    __shared__ uint32_t s_base[65];
    if (threadIdx.x < 64) {
        // lane l: class 64 - l; class 0 (empty tiles) last
        const uint32_t l = threadIdx.x;
        const uint32_t c = class_count[64 - l];
        const uint32_t inc = wave_incl_scan_u32(c);
        s_base[64 - l] = inc - c;
        if (l == 63) s_base[0] = inc;
        class_count[64 - l] = 0;
        if (l == 63) class_count[0] = 0;
    }
    block_scan_small<kSmallScanThreads>(buf, n);  // (its barriers order s_base)
    for (uint32_t i = threadIdx.x; i < n; i += kSmallScanThreads) {
        const uint32_t v = tile_cls[i];
        order[s_base[v >> 24] + (v & 0xffffffu)] = i;
    }
}

}  // namespace

int bin_emit(sfk_ctx* c, const BinJob& j) {
    hipStream_t st = c->stream;
    static_assert(kBinMaxCap < kSat, "a saturated isect total must exceed every capacity");
    if (j.n == 0 || j.cap < 2 || j.cap > kBinMaxCap || j.supers == 0 || j.supers >= (1u << 24) ||
        j.nviews == 0 || j.nviews > 255 || j.view_shift == 0 || j.view_shift > 10 ||
        j.supers % j.nviews || !j.counters || !j.status || !j.sstart ||
        (j.pairs_only ? !j.pairs_out : (!j.gids || !j.tile_starts)) ||
        (j.skip_place && !j.pairs_out))
        return fail(c, SFK_E_ARG, "binned emission: bad job");
    uint32_t* pair_keys = nullptr;
    uint2* pair_vals = nullptr;
    if (radix_scratch_spare_u32x2(j.sort, &pair_keys, &pair_vals))
        return fail(c, SFK_E_ARG, "binned emission: pair sort scratch");
    if ((j.seg_len ? j.seg_len : j.n) > 0x7fffffffu || !j.offs_by_splat)
        return fail(c, SFK_E_ARG, "binned emission: splat ids need 31 bits");
    if (j.pairs_only && j.pair_table && j.supers / j.nviews <= kPairMaxDigits)
        return bin_pairs_direct(c, j, j.supers / j.nviews);
    if (!j.status_zeroed)
        SFK_HIP(c, launch_fill_u32(st, j.status, 0u, scan_emit_status_words(j.n)));
    {
        SFK_PROF(c, "bin_scan");
        BinScanArgs a;
        a.ordered = j.ordered;
        a.n = j.n;
        a.seg_len = j.seg_len;
        a.views = j.views;
        a.view0 = j.view0;
        a.cap = j.cap;
        a.order_out = j.order_out;
        a.offsets = j.offsets;
        a.offs_by_splat = j.offs_by_splat;
        a.pair_keys = pair_keys;
        a.pair_vals = pair_vals;
        a.counters = j.counters;
        a.host_total = j.host_total;
        a.ticket = j.status;
        a.status = reinterpret_cast<uint64_t*>(j.status + 2);
        k_bin_scan<<<div_up(j.n, kRanks), kThreads, 0, st>>>(a);
        SFK_LAUNCHED(c, "bin scan");
    }
    // Pairs by supertile, stably (at least one bit, so the sorted pairs never
    // alias the input).
    uint32_t* sk = nullptr;
    uint2* sv = nullptr;
    int rc = argsort_u32_vals8(c, j.sort, pair_keys, pair_vals, j.cap, j.counters + 1,
                               j.view_shift, &sk, &sv);
    if (rc) return rc;
    // One view of at most 256 supertiles sorted in one pass: the pass's digit
    // totals are the supertiles' pair counts, and the count pass derives the
    // supertile ranges and chunk slots from them (no ranges launch).
    if (j.pairs_only) {
        SFK_PROF(c, "bin_ranges");
        SFK_HIP(c, launch_tile_starts(st, sk, j.cap, j.counters + 1, j.supers, j.sstart,
                                      j.nviews > 1 ? j.nviews : 0u, j.view_shift));
        *j.pairs_out = sv;
        return SFK_OK;
    }
    uint32_t db = 0;
    const uint32_t* totals = radix_scratch_totals(j.sort, &db);
    const bool ranges_in_count = j.nviews == 1 && j.view_shift <= db && j.supers <= 256;
    if (!ranges_in_count) {
        SFK_PROF(c, "bin_ranges");
        {
            SFK_HIP(c, launch_tile_starts(st, sk, j.cap, j.counters + 1, j.supers, j.sstart,
                                          j.nviews > 1 ? j.nviews : 0u, j.view_shift));
            k_bin_chunk_map<<<j.supers, 256, 0, st>>>(
                j.sstart, j.supers, j.views, j.nviews, reinterpret_cast<uint4*>(j.chunk_map));
            SFK_LAUNCHED(c, "bin chunk map");
        }
    }
    ChunkArgs ca;
    ca.pvals = sv;
    ca.sstart = j.sstart;
    ca.map = reinterpret_cast<const uint4*>(j.chunk_map);
    ca.offs_by_splat = j.offs_by_splat;
    ca.views = j.views;
    ca.view0 = j.view0;
    ca.supers = j.supers;
    ca.nviews = j.nviews;
    ca.cap = j.cap;
    ca.counts = j.chunk_counts;
    ca.tile_starts = j.tile_starts;
    ca.gids = j.gids;
    ca.tile_ids = j.tile_ids;
    ca.tile_order = nullptr;
    ca.order_tiles = 0;
    ca.totals = ranges_in_count ? totals : nullptr;
    ca.sstart_out = j.sstart;
    ca.map_out = reinterpret_cast<uint4*>(j.chunk_map);
    // With self_clean the count pass clears the emission scan's ticket and
    // look-back words (2 + 2 per workgroup of kRanks) and j.clear.
    ca.zero_a = j.self_clean ? j.status : nullptr;
    ca.zero_a_n = j.self_clean ? 2u + 2u * div_up(j.n, kRanks) : 0u;
    ca.zero_b = j.self_clean ? j.clear : nullptr;
    ca.zero_b_n = j.self_clean && j.clear ? j.clear_n : 0u;
    const uint32_t grid = std::min(div_up(bin_chunk_slots(j.cap, j.supers), 4), 2048u);
    {
        SFK_PROF(c, "bin_count");
        k_bin_chunks<false><<<grid, 256, 0, st>>>(ca);
        SFK_LAUNCHED(c, "bin count");
    }
    ca.totals = nullptr;
    ca.zero_a = ca.zero_b = nullptr;
    ca.zero_a_n = ca.zero_b_n = 0u;
    {
        SFK_PROF(c, "bin_tile_scan");
        if (j.zero_starts) SFK_HIP(c, launch_fill_u32(st, j.tile_starts, 0u, j.out_tiles));
        // Without a placement pass, the chunk scan counts the tile classes
        // and the tile scan writes the heaviest-first order.
        const bool order_here = j.skip_place && j.tile_order && j.class_count &&
                                j.out_tiles <= kSmallScanMax && j.out_tiles <= kTileOrderMax;
        k_bin_chunk_scan<<<j.supers, 256, 0, st>>>(j.sstart, j.views, j.view0, j.nviews,
                                                    j.chunk_counts, j.tile_starts,
                                                    order_here ? j.class_count : nullptr,
                                                    j.tile_cls);
        SFK_LAUNCHED(c, "bin chunk scan");
        if (order_here) {
            k_scan_small_order<<<1, kSmallScanThreads, 0, st>>>(j.tile_starts, j.out_tiles,
                                                                 j.tile_order, j.class_count,
                                                                 j.tile_cls);
            SFK_LAUNCHED(c, "bin tile scan");
        } else if (j.out_tiles <= kSmallScanMax) {
            k_scan_small<<<1, kSmallScanThreads, 0, st>>>(j.tile_starts, j.out_tiles);
            SFK_LAUNCHED(c, "bin tile scan");
        } else {
            SFK_HIP(c, launch_scan(st, nullptr, j.tile_starts, j.tile_starts, j.out_tiles,
                                   j.scan_sums, 0, j.tile_starts + j.out_tiles));
        }
    }
    if (j.skip_place) {
        // The rasterizers take each tile's isects from the sorted pairs
        // themselves (and write back the ones they consume).
        *j.pairs_out = sv;
        return SFK_OK;
    }
    {
        SFK_PROF(c, "bin_place");
        const bool order = j.tile_order && j.out_tiles <= kTileOrderMax;
        ca.tile_order = order ? j.tile_order : nullptr;
        ca.order_tiles = j.out_tiles;
        k_bin_chunks<true><<<grid + (order ? 1u : 0u), 256, 0, st>>>(ca);
        SFK_LAUNCHED(c, "bin place");
    }
    return SFK_OK;
}

}  // namespace sfk
