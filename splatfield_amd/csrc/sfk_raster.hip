// Raster stage for gfx950: intersection emission, tile ranges and the
// front-to-back compositor (src/raster.rs), for one view or for a batch of
// views sharing one sorted isect list (global tile id = view tile base +
// local tile id).
#include "sfk_device.h"
#include "sfk_internal.h"
#include "sfk_math.h"
#include "sfk_raster_core.h"

namespace sfk {

namespace {

// map_isects (raster.rs:12-48): rank r emits its bbox tiles, row-major, at
// offsets[r]; writes past max_isects are dropped while the walk continues.
// Ranks whose splat emits nothing (culled or off-screen) exit at once.
// Batched: rank r belongs to view r / seg_len; order[] holds in-view splat
// indices, per-splat arrays are view-major, and tile ids get the view's
// tile base.
__global__ __launch_bounds__(256) void k_map_isects(const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ tile_counts,
                                                    const uint2* __restrict__ bbox,
                                                    const uint32_t* __restrict__ offsets,
                                                    uint32_t n, uint32_t tiles_per_row,
                                                    uint32_t max_isects,
                                                    uint32_t* __restrict__ tile_ids,
                                                    uint32_t* __restrict__ gaussian_ids,
                                                    uint32_t seg_len,
                                                    const BatchView* __restrict__ views) {
    uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= n) return;
    uint32_t g = order[r];
    size_t gi = g;
    uint32_t tpr = tiles_per_row, tbase = 0;
    if (seg_len) {
        uint32_t v = r / seg_len;
        gi = (size_t)v * seg_len + g;
        tpr = views[v].tiles_x;
        tbase = views[v].tile_base;
    }
    if (tile_counts[gi] == 0) return;
    uint2 bb = bbox[gi];
    uint32_t min_x = bb.x & 0xffffu, min_y = bb.x >> 16, max_x = bb.y & 0xffffu,
             max_y = bb.y >> 16;
    uint32_t slot = offsets[r];
    for (uint32_t ty = min_y; ty < max_y; ++ty) {
        for (uint32_t tx = min_x; tx < max_x; ++tx) {
            if (slot < max_isects) {
                tile_ids[slot] = tbase + tx + ty * tpr;
                gaussian_ids[slot] = g;
            }
            ++slot;
        }
    }
}

// starts[t] = lower_bound(tile_ids, t) for t in [0, num_tiles] (raster.rs:
// 52-65): isect k writes every t in (tile_ids[k-1], tile_ids[k]], and k = n
// closes (tile_ids[n-1], num_tiles].  Each t is written exactly once.  A
// thread covers 4 consecutive isects with one 16-byte load.
__global__ __launch_bounds__(256) void k_tile_starts(const uint32_t* __restrict__ tile_ids,
                                                     uint32_t n_max,
                                                     const uint32_t* __restrict__ d_n,
                                                     uint32_t num_tiles,
                                                     uint32_t* __restrict__ starts,
                                                     uint32_t nviews, uint32_t view_shift) {
    // Batched keys (view << view_shift | local) map to local * nviews + view.
    auto id = [&](uint32_t key) {
        return nviews ? (key & ((1u << view_shift) - 1u)) * nviews + (key >> view_shift) : key;
    };
    uint32_t n = n_max;
    if (d_n) n = *d_n < n_max ? *d_n : n_max;
    // Grid-stride: a capacity-sized range runs on a bounded grid.
    for (uint64_t k00 = (uint64_t)(blockIdx.x * 256u + threadIdx.x) * 4u; k00 <= n;
         k00 += (uint64_t)gridDim.x * 1024u) {
        const uint32_t k0 = (uint32_t)k00;
        uint32_t ids[4];
        if (k0 + 4 <= n) {
            const uint4 v = reinterpret_cast<const uint4*>(tile_ids)[k0 / 4];
            ids[0] = id(v.x); ids[1] = id(v.y); ids[2] = id(v.z); ids[3] = id(v.w);
        } else {
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) ids[i] = k0 + i < n ? id(tile_ids[k0 + i]) : num_tiles;
        }
        uint32_t lo = k0 == 0 ? 0u : id(tile_ids[k0 - 1]) + 1u;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t k = k0 + i;
            if (k > n) break;
            const uint32_t hi = ids[i];  // num_tiles at k == n
            for (uint32_t t = lo; t <= hi; ++t) starts[t] = k;
            lo = hi + 1u;
        }
    }
}

// One tile of one view.
struct TileJob {
    uint32_t w, h, tx, ty, start, end, row_stride;
    float scale;
    const float4* proj;  // row base of this view
    uint32_t rows4;      // float4s per projected row (3 with colour, 2 without)
    // Split rows (batched): rows4 == 1, conic c from proj_c[g] and the
    // opacity from opac[g - obase].
    const float* proj_c;
    const float* opac;
    uint32_t obase;
    uint32_t* resp;
    uint32_t* fg;
    uint32_t* bg;
    // Pair source: the tile's supertile pairs [p0, p1), view-major splat
    // slot = in-view splat + gid_add.
    uint32_t p0, p1, gid_add;
};

// Pairs scanned per thread and step of stage_pairs: one (round, wave) hit
// mask per lane of the prefix scan, at most 16 rounds (32 VGPRs of pairs).
#ifndef SFK_PAIR_ROUNDS
#define SFK_PAIR_ROUNDS 4
#endif
template <int NT>
constexpr uint32_t pair_rounds() {
    return 64 / (NT / 64) < SFK_PAIR_ROUNDS ? 64 / (NT / 64) : SFK_PAIR_ROUNDS;
}

// Pair source of the rasterizers (one view or a batch): stages the tile's next
// isects, at most NT, into s_gid in pair order (= depth-rank order: the
// sorted pairs of one supertile and view keep the emission's rank order),
// from its supertile's pairs [p, J.p1); advances p past the last one taken
// and returns their count (0 once the pairs are exhausted).  A pair is an
// isect of tile (tx, ty) iff its compact bbox covers the tile and, for a
// pair flagged as possibly truncated, its emission slot off + (ty - y0) *
// (x1 - x0) + (tx - x0) is below the capacity (raster.rs:12-48 drops the
// others) -- exactly the isects that map_isects + the stable tile sort put
// in the tile's range, in the same order.  A step scans pair_rounds * NT
// pairs (pair p + k * NT + t in round k); one barrier per step publishes the
// waves' hit masks, from which every lane derives its hits' positions (a
// one-wavefront tile ranks them by running popcounts, with no barrier).
// Workgroup-uniform; ends with a barrier after the last write to s_gid.
// The pairs of one step of stage_pairs starting at q < pend.  Every lane
// issues all R loads (indices past the end clamped to the last pair, their
// bboxes cleared when used), so that the loads in flight are the same on
// every path and a prefetch is not drained early.
template <int NT>
__device__ __forceinline__ void pair_load(const RasterArgs& a, uint32_t pend, uint32_t q,
                                          uint2 (&pv)[pair_rounds<NT>()]) {
#pragma unroll
    for (uint32_t k = 0; k < pair_rounds<NT>(); ++k) {
        const uint32_t i = q + k * NT + threadIdx.x;
        pv[k] = a.pairs[i < pend ? i : pend - 1u];
    }
}

// `first` holds the pairs of the first step (pair_load at p, when p < J.p1),
// loaded by the caller ahead of its barrier.
template <int NT>
__device__ __forceinline__ uint32_t stage_pairs(const TileJob& J, const RasterArgs& a, uint32_t& p,
                                                uint2 (&first)[pair_rounds<NT>()],
                                                uint32_t* s_gid, uint64_t* s_mask,
                                                uint32_t* s_pnext) {
    constexpr uint32_t NW = NT / 64, R = pair_rounds<NT>();
    static_assert(R * NW <= 64, "one lane per (round, wave) mask");
    const uint32_t t = threadIdx.x, w = t / 64, lane = lane_id();
    const uint32_t pend = J.p1;
    auto load = [&](uint32_t q, uint2 (&pv)[R]) { pair_load<NT>(a, pend, q, pv); };
    uint32_t staged = 0, buf = 0;
    // One step over cur[] (loaded), issuing the next step's loads into nxt[]
    // while it ranks (that step starts at p + R * NT unless this one fills
    // the chunk, which ends the staging).  Returns whether to go on.
    auto step = [&](uint2 (&cur)[R], uint2 (&nxt)[R]) -> bool {
        bool hit[R];
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) {
            if (p + k * NT + t >= pend) cur[k].y = 0u;  // bbox 0 covers nothing
            const uint32_t cb = cur[k].y;
            const uint32_t x0 = cb & 0xffu, y0 = (cb >> 8) & 0xffu, x1 = (cb >> 16) & 0xffu,
                           y1 = cb >> 24;
            // x0 <= tx < x1 and y0 <= ty < y1 (x1 >= x0, y1 >= y0)
            bool h = (J.tx - x0 < x1 - x0) & (J.ty - y0 < y1 - y0);
            if (h & ((cur[k].x >> 31) != 0u)) {
                const uint32_t off = a.offs_by_splat[J.gid_add + (cur[k].x & 0x7fffffffu)];
                const uint32_t idx = (J.ty - y0) * (x1 - x0) + (J.tx - x0);
                h = off < a.cap && idx < a.cap - off;
            }
            hit[k] = h;
        }
        const uint32_t pn = p + R * NT;
        load(pn < pend ? pn : p, nxt);
        uint64_t m[R];
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) m[k] = __ballot(hit[k]);
        if constexpr (NW == 1) {
            // One wavefront: the hit masks are scalar and the positions are
            // running popcounts, with no exchange through LDS.
            uint32_t run = staged;
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) run += (uint32_t)__popcll(m[k]);
            uint32_t base = staged, resume = 0u;
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) {
                const uint32_t pos = base + mask_rank(m[k]);
                const bool take = hit[k];
                if (take && pos < (uint32_t)NT)
                    s_gid[pos] = (cur[k].x & 0x7fffffffu) + J.gid_add;
                if (run > (uint32_t)NT) {
                    // The last pair taken is the one at position NT - 1.
                    const uint64_t at = __ballot(take && pos == (uint32_t)NT - 1u);
                    if (at) resume = p + k * NT + (uint32_t)__builtin_ctzll(at) + 1u;
                }
                base += (uint32_t)__popcll(m[k]);
            }
            if (run > (uint32_t)NT) {
                p = resume;
                staged = NT;
                return false;
            }
            p = pn;
            staged = run;
            return staged < (uint32_t)NT && p < pend;
        }
        uint64_t* sm = s_mask + buf * (R * NW);
        buf ^= 1u;
        if (lane == 0) {
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) sm[k * NW + w] = m[k];
        }
        lds_barrier();
        // Lane j < R * NW of every wave: exclusive prefix of the hit counts
        // in (round, wave) order, i.e. in pair order.
        const uint32_t c = lane < R * NW ? (uint32_t)__popcll(sm[lane]) : 0u;
        const uint32_t inc = wave_incl_scan_u32(c);
        const uint32_t total = __builtin_amdgcn_readlane(inc, 63);
        const uint32_t run = staged + total;
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) {
            const uint32_t j = k * NW + w;
            const uint32_t base = staged + __shfl(inc - c, j, 64);
            if ((m[k] >> lane) & 1u) {
                const uint32_t pos = base + mask_rank(m[k]);
                if (pos < (uint32_t)NT) s_gid[pos] = (cur[k].x & 0x7fffffffu) + J.gid_add;
                if (pos == (uint32_t)NT - 1u && run > (uint32_t)NT) *s_pnext = p + k * NT + t + 1u;
            }
        }
        if (run > (uint32_t)NT) {
            // The last pair taken is the one at position NT - 1: resume after it.
            lds_barrier();
            p = *s_pnext;
            staged = NT;
            return false;
        }
        p = pn;
        staged = run;
        return staged < (uint32_t)NT && p < pend;
    };
    uint2 pong[R];
    if (p < pend) {
        while (step(first, pong) && step(pong, first)) {
        }
    }
    lds_barrier();
    return staged;
}

// rasterize_kernel (raster.rs:87-230) for one tile: TW x TW threads, one per
// pixel; splat rows staged through LDS in chunks of TW*TW.  Each wavefront
// walks only the staged splats whose reach box meets its 8 x 8 pixel block
// (the others cannot reach alpha >= 1/255 on any of its pixels, so skipping
// them changes nothing).  The responsibility modes add each splat's per-pixel
// u32 weights across the wavefront, then across the workgroup in LDS, and
// issue one global atomic per (splat, tile): the integer sum is order
// independent, so the result is bit-identical to one atomic per pixel.
// PF loads the staged rows of a group of 8 before compositing them, which
// overlaps their LDS latency at the cost of registers: it pays when the
// launch has too few waves to hide latency (one view's responsibility render)
// and costs occupancy otherwise.
// Barrier, then whether pred holds on any thread of the workgroup (a ballot
// when the workgroup is one wavefront).
template <int NW>
__device__ __forceinline__ bool block_any(bool pred) {
    if constexpr (NW == 1) {
        __syncthreads();
        return __ballot(pred) != 0u;
    } else {
        return __syncthreads_count(pred) != 0;
    }
}

// PIX: the per-pixel outputs (final T, n_contrib, colour) may be written;
// without them (the batched responsibility raster) no pixel tracks its last
// contributor.
template <int MODE, int TW, bool PF, bool PAIRS = false, bool PIX = true>
__device__ __forceinline__ void raster_tile(const TileJob& J, const RasterArgs& a) {

    constexpr int NT = TW * TW;
    constexpr int NW = NT / 64;
    __shared__ float4 s_geo[NT];  // mx, my, conic a, conic b
    __shared__ float4 s_aux[NT];  // conic c, opacity, r, g
    __shared__ float4 s_box[NT];  // reach box
    __shared__ float s_tau[NT];   // power bound of the ellipse test
    __shared__ uint16_t s_list[NW][NT];
    __shared__ float s_blue[MODE == kRgb ? NT : 1];
    __shared__ uint32_t s_gid[MODE == kRgb && !PAIRS ? 1 : NT];
    __shared__ uint32_t s_q0[MODE == kRgb ? 1 : NT];
    __shared__ uint32_t s_q1[MODE == kEvidence ? NT : 1];
    __shared__ uint64_t s_mask[PAIRS ? 2 * pair_rounds<NT>() * NW : 1];
    __shared__ uint32_t s_pnext;

    const uint32_t t = threadIdx.x;
    uint32_t lx, ly;
    tile_pixel<TW>(t, lx, ly);
    const uint32_t px = J.tx * TW + lx, py = J.ty * TW + ly;
    const bool in_bounds = px < J.w && py < J.h;
    const uint32_t start = J.start, end = J.end;
    const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
    const uint32_t wv = t / 64;
    const float bx0 = (float)(J.tx * TW + (lx & ~7u)), by0 = (float)(J.ty * TW + (ly & ~7u));
    uint16_t* list = s_list[wv];

    uint32_t inside = 0;
    if (MODE == kEvidence && in_bounds)
        inside = (a.mask[px / 32 + py * a.words_per_row] >> (px % 32)) & 1u;

    // The responsibility modes carry T scaled by their fixed-point scale (a
    // power of two, so the scaling is exact and commutes with every rounding
    // below): the weight `(alpha T) scale` is then one product alpha Ts, and
    // the saturation test T < 1/255 reads Ts < scale / 255.
    const float tscale = MODE == kRgb ? 1.0f : J.scale;
    const float tdone = kInv255 * tscale;
    float T = tscale, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !in_bounds;
    uint32_t last = start;

    uint32_t pnext = J.p0;
    for (uint32_t chunk = start; PAIRS || chunk < end; chunk += NT) {
        uint32_t cnt;
        if constexpr (PAIRS) {
            // The first step's pairs load across the barrier: previous chunk
            // fully consumed; whole-tile early exit before scanning further.
            uint2 first[pair_rounds<NT>()];
            if (pnext < J.p1) pair_load<NT>(a, J.p1, pnext, first);
            if (!block_any<NW>(!done)) break;
            cnt = stage_pairs<NT>(J, a, pnext, first, s_gid, s_mask, &s_pnext);
            if (cnt == 0) break;
            // Every chunk but the last holds NT isects, so `chunk` is the
            // slot of this one's first in the tile's range.
            if (a.gids_out && t < cnt) a.gids_out[chunk + t] = s_gid[t];
        } else {
            __syncthreads();  // previous chunk fully consumed
            cnt = (end - chunk) < (uint32_t)NT ? (end - chunk) : (uint32_t)NT;
        }
        const uint32_t idx = chunk + t;
        if (PAIRS ? t < cnt : idx < end) {
            uint32_t g = PAIRS ? s_gid[t] : a.gids[idx];
            const float4* row = J.proj + (size_t)g * J.rows4;
            float4 geo, aux;
            if (J.rows4 == 1) {
                geo = row[0];
                aux = make_float4(J.proj_c[g], J.opac[g - J.obase], 0.0f, 0.0f);
            } else {
                geo = row[0];
                aux = row[1];
            }
            s_box[t] = reach_box(geo, aux, &s_tau[t]);
            // Conic a and c staged halved, as the adjoint stages them.
            s_geo[t] = make_float4(geo.x, geo.y, 0.5f * geo.z, geo.w);
            s_aux[t] = make_float4(0.5f * aux.x, aux.y, aux.z, aux.w);
            if constexpr (MODE == kRgb) {
                s_blue[t] = row[2].x;
            } else {
                s_gid[t] = g;
                s_q0[t] = 0;
                if constexpr (MODE == kEvidence) s_q1[t] = 0;
            }
        }
        // Staging fence and whole-tile early exit (raster.rs:166-171).
        if (!block_any<NW>(!done)) break;
        const uint32_t nl =
            __ballot(!done) ? wave_reach_list<2>(s_box, s_tau, s_geo, s_aux, cnt, bx0, by0, list)
                            : 0u;
        // Listed splats in groups of 8 (entries past the list end repeat the
        // last entry and are masked off).  Each member is evaluated straight
        // through for the whole wave once any lane may be reached; lanes that
        // do not contribute keep their state through selects.  Per lane this
        // is raster.rs:183-217 exactly: alpha = min(o * exp(-power), 0.999),
        // contribute iff alpha >= 1/255, T *= 1 - alpha, done iff T < 1/255.
        const uint32_t lane = lane_id();
        const uint32_t ms = lane >> 3;
        for (uint32_t i0 = 0; i0 < nl; i0 += 8) {
            if (!__ballot(!done)) break;
            uint32_t jj[8];
            float4 geo[8], aux[8];
            float blue[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) jj[s] = list[min(i0 + s, nl - 1)];
            if constexpr (PF) {
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    geo[s] = s_geo[jj[s]];
                    aux[s] = s_aux[jj[s]];
                    if constexpr (MODE == kRgb) blue[s] = s_blue[jj[s]];
                    else blue[s] = 0.0f;
                }
            }
            uint32_t q[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) q[s] = 0;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                // RGB stops at the list end (wave-uniform); the responsibility
                // modes evaluate all 8 members straight through and mask the
                // ones past the end, so that the 8 weights need no merge of
                // control paths before their wave sum.
                if constexpr (MODE == kRgb) {
                    if (i0 + s >= nl) break;
                }
                const bool valid = MODE == kRgb || i0 + s < nl;
                float4 g4, x4;
                float bl = 0.0f;
                if constexpr (PF) {
                    g4 = geo[s]; x4 = aux[s]; bl = blue[s];
                } else {
                    g4 = s_geo[jj[s]]; x4 = s_aux[jj[s]];
                }
                const float dx = g4.x - pxf, dy = g4.y - pyf;
                float power;
                if constexpr (MODE == kRgb) {
                    // The adjoint's power and exp, operation for operation,
                    // so both passes take the same alpha and contributor
                    // decisions; the image is held to 1/255 of the exact
                    // evaluation, not to its bits.
                    power = fmaf(dx, fmaf(g4.z, dx, g4.w * dy), x4.x * dy * dy);
                } else {
                    // 0.5 (a dx dx + c dy dy) + b dx dy (raster.rs:183-186) with
                    // a and c staged halved: the halving is exact and
                    // commutes with the roundings, so the bits are the same.
                    power = (g4.z * dx * dx + x4.x * dy * dy) + g4.w * dx * dy;
                }
                // Every listed splat is in reach of some pixel of the block, so
                // only lanes that are done can leave a member unused: each
                // member is evaluated straight through (a wave-wide test
                // costs more than it saves).
                const bool in = valid && !done && !(power > kPowerSkip);
                // RGB: exp(-power) by the hardware exp2 (argument capped as in
                // the adjoint); the responsibility modes keep the exact exp
                // that their bit-exact u32 weights are specified with.
                const float G = MODE == kRgb
                                    ? __builtin_amdgcn_exp2f(fminf(-power * 1.44269504f, 126.0f))
                                    : sfk_expf_neg_in(power);
                const float alpha = fminf(x4.y * G, 0.999f);
                const bool con = in && alpha >= kInv255;
                const float vis = alpha * T;
                if constexpr (MODE == kRgb) {
                    if constexpr (!PF) bl = s_blue[jj[s]];
                    cr = con ? fmaf(x4.z, vis, cr) : cr;
                    cg = con ? fmaf(x4.w, vis, cg) : cg;
                    cb = con ? fmaf(bl, vis, cb) : cb;
                } else {
                    // `f32 as u32` of the weight (alpha T) scale = alpha Ts: on
                    // a contributing lane it lies in (0, 0.999 * 2^32], where
                    // the conversion truncates exactly as the saturating one
                    // would.
                    q[s] = con ? (uint32_t)vis : 0u;
                }
                T = con ? T * (1.0f - alpha) : T;
                if constexpr (PIX) last = con ? chunk + jj[s] + 1 : last;
                done = done || (con && T < tdone);
            }
            if constexpr (MODE != kRgb) {
                // The u32 weights of the 8 members are summed over the wave
                // in one transposed pass, lane l ending with member l >> 3.
                const uint32_t qor = (q[0] | q[1] | q[2] | q[3]) | (q[4] | q[5] | q[6] | q[7]);
                if (!__ballot(qor != 0u)) continue;
                const bool writer = (lane & 7u) == 0 && i0 + ms < nl;
                const uint32_t j = writer ? list[i0 + ms] : 0u;
                if constexpr (MODE == kTotal) {
                    uint32_t r = wave_sum8_transposed(q);
                    if (writer && r) atomicAdd(&s_q0[j], r);
                } else {
                    uint32_t qf[8], qb[8];
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        qf[s] = inside ? q[s] : 0u;
                        qb[s] = inside ? 0u : q[s];
                    }
                    uint32_t rf = wave_sum8_transposed(qf);
                    uint32_t rb = wave_sum8_transposed(qb);
                    if (writer) {
                        if (rf) atomicAdd(&s_q0[j], rf);
                        if (rb) atomicAdd(&s_q1[j], rb);
                    }
                }
            }
        }
        if constexpr (MODE != kRgb) {
            __syncthreads();
            if (t < cnt) {
                uint32_t g = s_gid[t];
                if constexpr (MODE == kTotal) {
                    if (s_q0[t]) atomicAdd(&J.resp[g], s_q0[t]);
                } else {
                    if (s_q0[t]) atomicAdd(&J.fg[g], s_q0[t]);
                    if (s_q1[t]) atomicAdd(&J.bg[g], s_q1[t]);
                }
            }
        }
    }

    if (PIX && in_bounds) {
        const size_t p = (size_t)py * J.w + px;
        if (a.final_t) a.final_t[p] = MODE == kRgb ? T : T * (1.0f / tscale);
        if (a.n_contrib) a.n_contrib[p] = last;
        if constexpr (MODE == kRgb) {
            if (a.rgba) reinterpret_cast<float4*>(a.rgba)[p] = make_float4(cr, cg, cb, 1.0f - T);
            if (a.bitmap) {
                uint32_t r = sfk_f2u(sfk_clampf(cr * 255.0f, 0.0f, 255.0f));
                uint32_t g = sfk_f2u(sfk_clampf(cg * 255.0f, 0.0f, 255.0f));
                uint32_t b = sfk_f2u(sfk_clampf(cb * 255.0f, 0.0f, 255.0f));
                uint32_t al = sfk_f2u(sfk_clampf((1.0f - T) * 255.0f, 0.0f, 255.0f));
                a.bitmap[px + (size_t)py * J.row_stride] = r | (g << 8) | (b << 16) | (al << 24);
            }
        }
    }
}

template <int MODE, int TW, bool PAIRS>
__global__ __launch_bounds__(TW * TW) void k_rasterize(RasterArgs a) {
    const uint32_t ntiles = a.tiles_x * a.tiles_y;
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, ntiles);
    TileJob J;
    J.w = a.w; J.h = a.h;
    J.tx = tile % a.tiles_x;
    J.ty = tile / a.tiles_x;
    J.start = a.tile_starts[tile];
    J.end = a.tile_starts[tile + 1];
    J.row_stride = a.row_stride;
    J.scale = a.scale;
    J.proj = reinterpret_cast<const float4*>(a.proj);
    J.rows4 = 3;
    J.proj_c = nullptr;
    J.opac = nullptr;
    J.obase = 0;
    J.resp = a.resp; J.fg = a.fg; J.bg = a.bg;
    J.p0 = J.p1 = 0;
    J.gid_add = 0;
    if constexpr (PAIRS) {
        const uint32_t ls =
            (J.ty / kBinSuper) * ((a.tiles_x + kBinSuper - 1) / kBinSuper) + J.tx / kBinSuper;
        J.p0 = a.pair_starts[ls];
        J.p1 = a.pair_starts[ls + 1];
    }
    raster_tile<MODE, TW, MODE != kRgb, PAIRS>(J, a);
}

// Batched responsibility: one launch over the tiles of every view, from the
// sorted isect lists or (PAIRS) straight from the supertile pairs.
template <int TW, bool PAIRS>
__global__ __launch_bounds__(TW * TW) void k_rasterize_batch(RasterArgs a) {
    const uint32_t total = a.batch_tiles;
    uint32_t lo, local;
    if (a.tile_order) {
        // A start-table slot local * nviews + view in launch order
        // (heaviest first, or dealt to the XCDs); slots past a view's grid
        // (smaller views, idle workgroups) are skipped.
        const uint32_t slot = a.tile_order[blockIdx.x];
        lo = slot % a.nviews;
        local = slot / a.nviews;
        if (local >= a.views[lo].tiles_x * a.views[lo].tiles_y) return;
    } else {
        const uint32_t tile = xcd_remap(blockIdx.x, total);
        // view of this tile: last view whose tile base <= tile
        lo = 0;
        uint32_t hi = a.nviews;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (a.views[mid].tile_base <= tile) lo = mid;
            else hi = mid;
        }
        local = tile - a.views[lo].tile_base;
    }
    const BatchView bv = a.views[lo];
    TileJob J;
    J.w = bv.w; J.h = bv.h;
    J.tx = local % bv.tiles_x;
    J.ty = local / bv.tiles_x;
    // Sorted by (local tile, view): list (view lo, tile local) is entry
    // local * nviews + lo of the starts table; ids are view-major slots.
    if constexpr (PAIRS) {
        const uint32_t sh = a.super_shift;
        const uint32_t ls = (J.ty >> sh) * ((bv.tiles_x + (1u << sh) - 1u) >> sh) + (J.tx >> sh);
        const uint32_t ps = ls * a.nviews + lo;
        J.p0 = a.pair_starts[ps];
        J.p1 = a.pair_starts[ps + 1];
        J.start = 0;
        J.end = 0;
    } else {
        const uint32_t slot = local * a.nviews + lo;
        J.start = a.tile_starts[slot];
        J.end = a.tile_starts[slot + 1];
        J.p0 = J.p1 = 0;
    }
    J.gid_add = lo * a.seg_len;
    J.row_stride = 0;
    J.scale = bv.scale;
    J.proj = reinterpret_cast<const float4*>(a.proj);
    J.rows4 = 1;
    J.proj_c = a.proj_c;
    J.opac = a.opac;
    J.obase = lo * a.seg_len;  // ids are view-major: view lo's splat i is lo * n + i
    J.resp = a.resp;
    J.fg = nullptr; J.bg = nullptr;
    RasterArgs b = a;
    b.final_t = nullptr;
    b.n_contrib = nullptr;
    raster_tile<kTotal, TW, false, PAIRS, false>(J, b);
}

template <int MODE, bool PAIRS>
hipError_t raster_tw(hipStream_t s, const RasterArgs& a) {
    uint32_t nt = a.tiles_x * a.tiles_y;
    if (nt == 0) return hipSuccess;
    switch (a.tile_w) {
        case 8: k_rasterize<MODE, 8, PAIRS><<<nt, 64, 0, s>>>(a); break;
        case 16: k_rasterize<MODE, 16, PAIRS><<<nt, 256, 0, s>>>(a); break;
        case 32: k_rasterize<MODE, 32, PAIRS><<<nt, 1024, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t raster_src(hipStream_t s, const RasterArgs& a) {
    return a.pairs ? raster_tw<MODE, true>(s, a) : raster_tw<MODE, false>(s, a);
}

}  // namespace

hipError_t launch_map_isects(hipStream_t s, const uint32_t* order, const uint32_t* tile_counts,
                             const uint32_t* bbox, const uint32_t* offsets, uint32_t n,
                             uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                             uint32_t* gaussian_ids, uint32_t seg_len, const BatchView* views) {
    if (n == 0) return hipSuccess;
    k_map_isects<<<div_up(n, 256), 256, 0, s>>>(order, tile_counts,
                                                reinterpret_cast<const uint2*>(bbox), offsets, n,
                                                tiles_per_row, max_isects, tile_ids, gaussian_ids,
                                                seg_len, views);
    return hipGetLastError();
}

hipError_t launch_tile_starts(hipStream_t s, const uint32_t* tile_ids, uint32_t n,
                              const uint32_t* d_n, uint32_t num_tiles, uint32_t* starts,
                              uint32_t nviews, uint32_t view_shift) {
    uint32_t grid = div_up((uint64_t)n + 1, 1024);
    if (d_n && grid > 4096) grid = 4096;  // the live count may be far below n
    k_tile_starts<<<grid, 256, 0, s>>>(tile_ids, n, d_n, num_tiles, starts, nviews, view_shift);
    return hipGetLastError();
}

hipError_t launch_rasterize(hipStream_t s, int mode, const RasterArgs& a) {
    switch (mode) {
        case kRgb: return raster_src<kRgb>(s, a);
        case kTotal: return raster_src<kTotal>(s, a);
        case kEvidence: return raster_src<kEvidence>(s, a);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rasterize_batch(hipStream_t s, const RasterArgs& a) {
    if (a.batch_tiles == 0) return hipSuccess;
    if (a.pairs) {
        switch (a.tile_w) {
            case 8: k_rasterize_batch<8, true><<<a.batch_tiles, 64, 0, s>>>(a); break;
            case 16: k_rasterize_batch<16, true><<<a.batch_tiles, 256, 0, s>>>(a); break;
            case 32: k_rasterize_batch<32, true><<<a.batch_tiles, 1024, 0, s>>>(a); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (a.tile_w) {
            case 8: k_rasterize_batch<8, false><<<a.batch_tiles, 64, 0, s>>>(a); break;
            case 16: k_rasterize_batch<16, false><<<a.batch_tiles, 256, 0, s>>>(a); break;
            case 32: k_rasterize_batch<32, false><<<a.batch_tiles, 1024, 0, s>>>(a); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace sfk
