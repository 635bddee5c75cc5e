// Batched EIG round for gfx950: every candidate view of one B3-Seg round
// (src/seg/active.rs:113-129 scores them one at a time, each with two host
// syncs) as one stream-ordered sequence, optionally captured once into a
// hipGraph and replayed per round:
//
//   project all views (attributes read once per splat, geometry only: the
//   responsibility does not depend on colour) -> one segmented depth argsort
//   (one segment per view) -> one gather scan over all views -> isect
//   emission with global tile ids (view tile base + local tile) -> one stable
//   argsort of all isects by global tile id (= view|tile|depth order) ->
//   tile ranges -> one responsibility raster over every view's tiles, each
//   view at its own fixed-point scale -> per-view EIG with f64 entropy ->
//   device argmax.
//
// The per-pixel and per-splat arithmetic is the single-view code's
// (sfk_project_core.h, raster_tile), so each view's u32 evidence equals a
// plain render of that view bit for bit.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "sfk_entropy.h"
#include "sfk_project_core.h"

using namespace sfk;

namespace sfk {
namespace {

constexpr uint32_t kMaxViews = 64;

// One thread per splat, looping over the views: the 44-byte attribute row is
// read once for all views.  The projected rows are split: per (view, splat)
// proj = (mean x, mean y, conic a, conic b) and proj_c = conic c; the
// opacity, the same in every view, once per splat (opac).  It also clears
// the round's responsibility words (resp, per view and splat), which the
// raster then accumulates into.
__global__ __launch_bounds__(256) void k_project_batch(const sfk_view* __restrict__ views,
                                                       uint32_t nviews,
                                                       const float* __restrict__ A, uint32_t n,
                                                       float inv_tile,
                                                       uint32_t* __restrict__ depth_keys,
                                                       float4* __restrict__ proj,
                                                       float* __restrict__ proj_c,
                                                       float* __restrict__ opac,
                                                       uint32_t* __restrict__ tile_counts,
                                                       uint2* __restrict__ bbox,
                                                       uint32_t* __restrict__ cbbox,
                                                       uint32_t* __restrict__ counters,
                                                       uint32_t* __restrict__ resp) {
    __shared__ sfk_view s_views[kMaxViews];
    __shared__ float2 s_lims[kMaxViews];
    // Informational per-view counts (isects, visible): each wave adds its
    // sums here, the workgroup adds them to one of kCountSlots slot pairs
    // per view once, after the loop (no barrier per view).
    __shared__ uint32_t s_cnt[kMaxViews][2];
    for (uint32_t k = threadIdx.x; k < nviews * (sizeof(sfk_view) / 4); k += 256)
        reinterpret_cast<float*>(s_views)[k] = reinterpret_cast<const float*>(views)[k];
    for (uint32_t k = threadIdx.x; k < 2 * nviews; k += 256) (&s_cnt[0][0])[k] = 0u;
    __syncthreads();
    // Per view once, not per splat and view: the clamp limits.
    for (uint32_t k = threadIdx.x; k < nviews; k += 256) s_lims[k] = view_lims(s_views[k]);
    __syncthreads();
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool live = i < n;
    SplatIn s;
    float m[9];
    float cutoff = 0.0f;
    if (live) {
        s = load_splat(A, i, n);
        opac[i] = s.opacity;
        // Per splat once, not per view: R S and the extent cutoff.
        splat_rs(s, m);
        cutoff = splat_cutoff(s);
    }
    for (uint32_t v = 0; v < nviews; ++v) {
        bool vis = false;
        uint32_t cnt = 0;
        if (live) {
            ProjGeom o;
            const size_t gi = (size_t)v * n + i;
            uint32_t key = 0xffffffffu;
            if (project_geom_pre(s_views[v], s, m, cutoff, s_lims[v], inv_tile, o)) {
                vis = true;
                key = __builtin_bit_cast(uint32_t, o.cz);
                cnt = o.count;
                proj[gi] = make_float4(o.px, o.py, o.qa, o.qb);
                proj_c[gi] = o.qc;
                if (!cbbox) bbox[gi] = make_uint2(o.bb_lo, o.bb_hi);
            }
            depth_keys[gi] = key;
            resp[gi] = 0u;
            // Compact grids: the bbox word is all the emission needs.
            if (cbbox) cbbox[gi] = vis ? compact_bbox(o) : 0u;
            else tile_counts[gi] = cnt;
        }
        const uint32_t wc = wave_sum_last_u32(cnt);  // valid in lane 63
        const uint32_t wvis = (uint32_t)__popcll(__ballot(vis));
        if (lane_id() == 63) {
            if (wc) atomicAdd(&s_cnt[v][0], wc);
            if (wvis) atomicAdd(&s_cnt[v][1], wvis);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 2 * nviews; k += 256) {
        const uint32_t c = (&s_cnt[0][0])[k];
        if (c) atomicAdd(&counters[2 + ((size_t)(k / 2) * kCountSlots + blockIdx.x % kCountSlots) * 2 +
                                   (k & 1u)], c);
    }
}

// H(a_i, b_i) once per round, shared by every view.
__global__ __launch_bounds__(256) void k_entropy_base(const float* __restrict__ a,
                                                      const float* __restrict__ b, uint32_t n,
                                                      double* __restrict__ h0) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) h0[i] = d_beta_entropy((double)a[i], (double)b[i]);
}

constexpr uint32_t kEigBlock = 1024;

// grid (blocks per view, views): per-block f64 partial sums of the entropy
// drops of one view.  Only splats with a nonzero responsibility have a drop,
// so each wavefront first compacts its 256 elements to the nonzero ones
// (ballot order, in LDS) and then evaluates them 64 at a time with every
// lane busy, instead of leaving the lanes of zero weights idle through the
// f64 entropy.  The summation order is fixed by the compacted order.
__global__ __launch_bounds__(256) void k_eig_batch(const float* __restrict__ a,
                                                   const float* __restrict__ b,
                                                   const double* __restrict__ h0,
                                                   const uint32_t* __restrict__ resp, uint32_t n,
                                                   const BatchView* __restrict__ views,
                                                   double* __restrict__ partials) {
    __shared__ double s_w[4];
    __shared__ uint32_t s_i[4][256], s_q[4][256];
    const uint32_t v = blockIdx.y;
    const float scale = views[v].scale;
    const uint32_t* r = resp + (size_t)v * n;
    const uint32_t w = threadIdx.x / 64, lane = lane_id();
    // Wave w owns elements base + [w * 256, (w + 1) * 256).
    const uint32_t base = blockIdx.x * kEigBlock + w * 256u;
    uint32_t q[4];
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t i = base + e * 64 + lane;
        q[e] = i < n ? r[i] : 0u;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint64_t m = __ballot(q[e] != 0u);
        if (q[e]) {
            const uint32_t p = cnt + mask_rank(m);
            s_i[w][p] = base + e * 64 + lane;
            s_q[w][p] = q[e];
        }
        cnt += (uint32_t)__popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    double sum = 0.0;
    for (uint32_t j = lane; j < cnt; j += 64) {
        const uint32_t i = s_i[w][j];
        sum += d_entropy_drop(a[i], b[i], h0[i], s_q[w][j], scale);
    }
    sum = wave_sum_f64(sum);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(size_t)v * gridDim.x + blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// One workgroup per view: fixed-order sum of its partials.
__global__ __launch_bounds__(256) void k_eig_batch_final(const double* __restrict__ partials,
                                                         uint32_t per_view,
                                                         double* __restrict__ eig) {
    __shared__ double s_w[4];
    const double* p = partials + (size_t)blockIdx.x * per_view;
    double sum = 0.0;
    for (uint32_t i = threadIdx.x; i < per_view; i += 256) sum += p[i];
    sum = wave_sum_f64(sum);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) eig[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// select_view's argmax (active.rs:116-127): strictly greater wins, so ties
// keep the first candidate.
__global__ void k_argmax(const double* __restrict__ eig, uint32_t nviews,
                         uint32_t* __restrict__ best) {
    if (threadIdx.x != 0) return;
    uint32_t bi = 0;
    double bv = -INFINITY;
    for (uint32_t v = 0; v < nviews; ++v)
        if (eig[v] > bv) {
            bv = eig[v];
            bi = v;
        }
    *best = bi;
}

}  // namespace
}  // namespace sfk

struct sfk_eig_round {
    sfk_ctx* ctx = nullptr;
    const sfk_splats* splats = nullptr;
    uint32_t n = 0, nviews = 0, tile_w = 16, digit_bits = 8;
    uint32_t total_tiles = 0, capacity = 0;
    uint32_t max_tiles = 0, tile_bits = 0;  // largest view's tiles, bits_for of it
    bool compact = true;                     // every view's grid fits the compact bbox
    uint32_t* cbbox = nullptr;
    std::vector<sfk_view> views;
    std::vector<BatchView> bviews;
    sfk_view* d_views = nullptr;
    BatchView* d_bviews = nullptr;
    float4* proj = nullptr;    // per (view, splat): mean x, y, conic a, b
    float* proj_c = nullptr;   // per (view, splat): conic c
    float* opac = nullptr;     // per splat: opacity
    uint32_t* keys = nullptr;
    uint32_t* counts = nullptr;
    uint2* bbox = nullptr;
    uint32_t* counters = nullptr;  // [0] total isects (scan); 2 + v*2*kCountSlots: slots
    uint32_t* tile_ids = nullptr;
    uint32_t* gids = nullptr;
    uint32_t* starts = nullptr;
    uint32_t* resp = nullptr;
    double* h0 = nullptr;
    double* partials = nullptr;
    double* eig = nullptr;
    uint32_t* best = nullptr;
    double* host_out = nullptr;  // pinned: eig[V], then best as a double slot
    sfk_radix_scratch* sort_depth = nullptr;
    sfk_radix_scratch* sort_tile = nullptr;
    sfk_scan_scratch* scan = nullptr;
    uint32_t* emit_status = nullptr;  // look-back words of the fused scan/emission
    const uint32_t* sorted_tiles = nullptr;
    const uint32_t* sorted_gids = nullptr;
    // Binned emission (compact grids): per-view supertile grids, supertile
    // ranges, chunk map and counts, per-rank isect offsets.
    bool binned = true;
    // Binned: the raster takes each tile's isects straight from its
    // supertile's sorted pairs (no per-tile lists are placed).
    bool pair_raster = true;
    const uint2* pairs = nullptr;
    // Scratch of the direct pair emission (views of at most kPairMaxDigits
    // supertiles).
    uint32_t* pair_table = nullptr;
    uint32_t* pair_table_sums = nullptr;
    uint32_t supers = 0, bin_shift = 1;  // supertile ids, bits of the local ones
    BinView* d_binviews = nullptr;
    // Views of the direct pair emission: supertiles of 2^pair_shift tiles a
    // side, at most pair_supers per view.
    BinView* d_pairviews = nullptr;
    uint32_t pair_shift = kBinSuperShift, pair_supers = 0;
    // Launch order of the pair raster (start-table slots local * views +
    // view, ~0u for an idle workgroup): workgroup b runs on XCD b % 8, and
    // every supertile row of every view goes whole to one XCD, dealt on a
    // diagonal so that each XCD gets a like share of every view.
    uint32_t* xcd_order = nullptr;
    uint32_t xcd_order_len = 0;
    uint32_t* offsets = nullptr;  // isect offsets of truncation candidates, by view-major splat
    uint32_t* bin_sstart = nullptr;
    uint32_t* bin_map = nullptr;
    uint32_t* bin_counts = nullptr;
    uint32_t* scan_sums = nullptr;
    // Tile-start slots heaviest first (the batched raster's launch order),
    // written by the binned emission when the slot table is small enough.
    uint32_t* tile_order = nullptr;
    // Depth keys rebased per view to 24 bits when every view's AABB depth
    // span allows it (depth_rebase).
    uint32_t depth_bits = 32;
    uint32_t* d_rebase = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const float* graph_a = nullptr;
    const float* graph_b = nullptr;
};

extern "C" void sfk_eig_round_destroy(sfk_eig_round* r);

static void round_free_isects(sfk_eig_round* r) {
    void* bufs[] = {r->tile_ids, r->gids, r->bin_map, r->bin_counts};
    for (void* p : bufs)
        if (p) hipFree(p);
    r->tile_ids = r->gids = r->bin_map = r->bin_counts = nullptr;
    sfk_radix_scratch_destroy(r->sort_tile);
    r->sort_tile = nullptr;
}

static int round_alloc_isects(sfk_eig_round* r, uint32_t cap) {
    sfk_ctx* c = r->ctx;
    round_free_isects(r);
    SFK_HIP(c, hipMalloc(&r->tile_ids, (size_t)cap * 4));
    SFK_HIP(c, hipMalloc(&r->gids, (size_t)cap * 4));
    if (r->compact && cap <= kBinMaxCap) {
        const size_t slots = bin_chunk_slots(cap, r->supers);
        SFK_HIP(c, hipMalloc(&r->bin_map, slots * 16));
        SFK_HIP(c, hipMalloc(&r->bin_counts, slots * 64 * 4));
    }
    // 8-byte values where the binned emission may sort its pairs.
    int rc = radix_scratch_create_ex(c, cap, 4, r->compact && cap <= kBinMaxCap ? 8 : 4,
                                     &r->sort_tile);
    if (rc) return rc;
    sfk_radix_set_digit_bits(r->sort_tile, r->digit_bits);
    r->capacity = cap;
    return SFK_OK;
}

// The whole round, enqueued on the context stream.
static int round_enqueue(sfk_eig_round* r, const float* a, const float* b) {
    sfk_ctx* c = r->ctx;
    const uint32_t n = r->n, V = r->nviews;
    const size_t vn = (size_t)V * n;
    const float* attrs = nullptr;
    sfk_splats_info(r->splats, nullptr, nullptr, nullptr, nullptr, &attrs, nullptr);
    const bool binned = r->compact && r->binned && r->capacity >= 2 && r->capacity <= kBinMaxCap;
    SFK_HIP(c, launch_zero2_u32(c->stream, r->counters, 2 + (size_t)V * 2 * kCountSlots,
                                r->emit_status, binned ? scan_emit_status_words((uint32_t)vn) : 0u));
    {
        SFK_PROF(c, "batch_project");
        k_project_batch<<<div_up(n, 256), 256, 0, c->stream>>>(
            r->d_views, V, attrs, n, 1.0f / (float)r->tile_w, r->keys, r->proj, r->proj_c,
            r->opac, r->counts,
            r->bbox, r->compact ? r->cbbox : nullptr, r->counters, r->resp);
        SFK_LAUNCHED(c, "batch project");
    }
    EmitSource src;
    int rc;
    bool emitted = false;
    if (r->compact) {
        uint2* ordered = nullptr;
        KeyRebase rb;
        rb.bases = r->d_rebase;
        rc = argsort_payload_u32(c, r->sort_depth, r->keys, r->cbbox, n, V, r->depth_bits,
                                 &ordered, r->d_rebase ? &rb : nullptr);
        if (rc) return rc;
        src.ordered = ordered;
        if (binned) {
            // Every (view, tile) list in depth order, no isect sort; tile
            // ranges indexed local * V + view as the batched raster reads them.
            BinJob j;
            memset(&j, 0, sizeof j);
            j.ordered = ordered;
            j.n = (uint32_t)vn;
            j.seg_len = n;
            j.views = r->d_binviews;
            j.nviews = V;
            j.view_shift = r->bin_shift;
            j.supers = r->supers;
            j.out_tiles = r->max_tiles * V;
            j.zero_starts = true;
            j.cap = r->capacity;
            j.offs_by_splat = r->offsets;
            j.sort = r->sort_tile;
            j.sstart = r->bin_sstart;
            j.chunk_map = r->bin_map;
            j.chunk_counts = r->bin_counts;
            j.scan_sums = r->scan_sums;
            j.tile_starts = r->starts;
            j.tile_order = r->tile_order;
            j.gids = r->gids;
            j.counters = r->counters;
            j.status = r->emit_status;
            j.status_zeroed = true;
            j.pairs_only = r->pair_raster;
            j.pairs_out = &r->pairs;
            j.pair_table = r->pair_table;
            j.pair_table_sums = r->pair_table_sums;
            if (r->pair_raster && r->pair_table) {
                // The direct pair emission, on its own supertile grids.
                j.views = r->d_pairviews;
                j.supers = r->pair_supers * V;
            }
            rc = bin_emit(c, j);
            if (rc) return rc;
            r->sorted_tiles = nullptr;
            r->sorted_gids = r->gids;
            emitted = true;
        }
    } else {
        uint32_t* order = nullptr;
        rc = sfk_argsort_u32_segmented(c, r->sort_depth, r->keys, nullptr, V, n, 32, 0, nullptr,
                                       &order);
        if (rc) return rc;
        src.order = order;
        src.tile_counts = r->counts;
        src.bbox = reinterpret_cast<const uint32_t*>(r->bbox);
    }
    if (!emitted) {
        {
            SFK_PROF(c, "scan_emit");
            SFK_HIP(c, launch_scan_emit(c->stream, src, (uint32_t)vn, 0, r->capacity,
                                        r->tile_ids, r->gids, nullptr, r->counters,
                                        r->emit_status, n, r->d_bviews, r->tile_bits));
        }
        uint32_t* st = nullptr;
        uint32_t* sg = nullptr;
        // Only the local tile bits are sorted: the emission is view-major, so
        // a stable sort on them orders by (local tile, view, depth rank).
        rc = sfk_argsort_u32(c, r->sort_tile, r->tile_ids, r->gids, r->capacity, r->counters,
                             r->tile_bits, 1, &st, &sg);
        if (rc) return rc;
        r->sorted_tiles = st;
        r->sorted_gids = sg;
        {
            SFK_PROF(c, "tile_starts");
            SFK_HIP(c, launch_tile_starts(c->stream, st, r->capacity, r->counters,
                                          r->max_tiles * V, r->starts, V, r->tile_bits));
        }
    }
    {
        SFK_PROF(c, "batch_raster_total");
        RasterArgs ra;
        memset(&ra, 0, sizeof ra);
        ra.tile_w = r->tile_w;
        ra.tile_starts = r->starts;
        ra.gids = r->sorted_gids;
        ra.proj = reinterpret_cast<const float*>(r->proj);
        ra.proj_c = r->proj_c;
        ra.opac = r->opac;
        ra.resp = r->resp;
        ra.views = r->d_bviews;
        ra.nviews = V;
        ra.batch_tiles = r->total_tiles;
        ra.seg_len = n;
        // (slot order of the binned emission; the sorted path writes none)
        ra.tile_order = r->sorted_tiles == nullptr ? r->tile_order : nullptr;
        if (binned && r->pair_raster) {
            ra.tile_order = nullptr;
            ra.pairs = r->pairs;
            ra.pair_starts = r->bin_sstart;
            ra.offs_by_splat = r->offsets;
            ra.cap = r->capacity;
            ra.super_shift = r->pair_table ? r->pair_shift : kBinSuperShift;
            if (r->pair_table && r->xcd_order) {
                ra.tile_order = r->xcd_order;
                ra.batch_tiles = r->xcd_order_len;
            }
        }
        SFK_HIP(c, launch_rasterize_batch(c->stream, ra));
    }
    {
        SFK_PROF(c, "batch_eig");
        const uint32_t bpv = div_up(n, kEigBlock);
        k_entropy_base<<<div_up(n, 256), 256, 0, c->stream>>>(a, b, n, r->h0);
        k_eig_batch<<<dim3(bpv, V), 256, 0, c->stream>>>(a, b, r->h0, r->resp, n, r->d_bviews,
                                                          r->partials);
        k_eig_batch_final<<<V, 256, 0, c->stream>>>(r->partials, bpv, r->eig);
        k_argmax<<<1, 64, 0, c->stream>>>(r->eig, V, r->best);
        SFK_LAUNCHED(c, "batch eig");
    }
    SFK_HIP(c, hipMemcpyAsync(r->host_out, r->eig, (size_t)V * 8, hipMemcpyDeviceToHost,
                              c->stream));
    SFK_HIP(c, hipMemcpyAsync(r->host_out + V, r->best, 4, hipMemcpyDeviceToHost, c->stream));
    return SFK_OK;
}

extern "C" int sfk_eig_round_create(sfk_ctx* c, const sfk_splats* sp, const sfk_camera* cams,
                                    const uint32_t* widths, const uint32_t* heights,
                                    uint32_t nviews, uint32_t tile_w, uint32_t digit_bits,
                                    sfk_eig_round** out) {
    if (!c || !sp || !cams || !widths || !heights || !out) return SFK_E_ARG;
    *out = nullptr;
    if (nviews == 0 || nviews > kMaxViews)
        return fail(c, SFK_E_ARG, "eig round: %u views (1..%u)", nviews, kMaxViews);
    if (!(tile_w == 8 || tile_w == 16 || tile_w == 32) || digit_bits < 4 || digit_bits > 8)
        return fail(c, SFK_E_ARG, "eig round: tile width %u / digit bits %u", tile_w, digit_bits);
    uint32_t n = 0;
    sfk_splats_info(sp, &n, nullptr, nullptr, nullptr, nullptr, nullptr);
    if ((uint64_t)n * nviews > 0xffffffffull)
        return fail(c, SFK_E_CAPACITY, "eig round: %u views x %u splats exceeds u32", nviews, n);
    auto* r = new (std::nothrow) sfk_eig_round();
    if (!r) return SFK_E_NOMEM;
    r->ctx = c;
    r->splats = sp;
    r->n = n;
    r->nviews = nviews;
    r->tile_w = tile_w;
    r->digit_bits = digit_bits;
    uint64_t tiles = 0, cap0 = 0;
    for (uint32_t v = 0; v < nviews; ++v) {
        sfk_view vw;
        int rc = sfk_camera_view(&cams[v], widths[v], heights[v], tile_w, &vw);
        if (rc) {
            delete r;
            return fail(c, SFK_E_ARG, "eig round: view %u rejected", v);
        }
        r->views.push_back(vw);
        BatchView bv;
        memset(&bv, 0, sizeof bv);
        bv.w = widths[v];
        bv.h = heights[v];
        bv.tiles_x = (widths[v] + tile_w - 1) / tile_w;
        bv.tiles_y = (heights[v] + tile_w - 1) / tile_w;
        bv.tile_base = (uint32_t)tiles;
        bv.scale = (float)(1ull << sfk_fractional_bits(widths[v], heights[v]));
        tiles += (uint64_t)bv.tiles_x * bv.tiles_y;
        r->max_tiles = std::max(r->max_tiles, bv.tiles_x * bv.tiles_y);
        if (bv.tiles_x > kCompactGrid || bv.tiles_y > kCompactGrid) r->compact = false;
        cap0 += std::min<uint64_t>((uint64_t)bv.tiles_x * bv.tiles_y * n, 1u << 22);
        r->bviews.push_back(bv);
    }
    r->total_tiles = (uint32_t)tiles;
    r->tile_bits = sfk_bits_for(r->max_tiles);
    // Supertile grids of the binned emission.  Ids are local * views + view:
    // the view-major pairs sorted stably on their local bits alone come out
    // in that order (one radix pass up to 256 supertiles per view).
    std::vector<BinView> binviews(nviews);
    uint32_t max_supers = 1;
    for (uint32_t v = 0; v < nviews; ++v)
        max_supers = std::max(max_supers, bin_supers(r->bviews[v].tiles_x, r->bviews[v].tiles_y));
    r->bin_shift = std::max(1u, sfk_bits_for(max_supers));
    r->supers = max_supers * nviews;
    for (uint32_t v = 0; v < nviews; ++v) {
        BinView& b = binviews[v];
        memset(&b, 0, sizeof b);
        b.tiles_x = r->bviews[v].tiles_x;
        b.tiles_y = r->bviews[v].tiles_y;
        b.supers_x = div_up(b.tiles_x, kBinSuper);
        b.key_base = v << r->bin_shift;
        b.out_mul = nviews;
        b.out_add = v;
        b.gid_add = v * n;
        b.super_shift = kBinSuperShift;
    }
    // The direct pair emission's supertiles: kPairSuperPx pixels a side, at
    // least 2 x 2 tiles.  Smaller supertiles mean more pairs per splat but a
    // shorter scan of its supertile's pairs for every tile.
    // The pair write ranks spans of at most kPairMaxSpan supertiles a row,
    // and the direct emission takes at most kPairMaxDigits a view: larger
    // views get larger supertiles.
    r->pair_shift = 1;
    while ((r->tile_w << r->pair_shift) < kPairSuperPx) ++r->pair_shift;
    for (uint32_t v = 0; v < nviews; ++v) {
        const uint32_t tx = r->bviews[v].tiles_x, ty = r->bviews[v].tiles_y;
        while (r->pair_shift < 8 &&
               (div_up(tx, 1u << r->pair_shift) > kPairMaxSpan ||
                div_up(tx, 1u << r->pair_shift) * div_up(ty, 1u << r->pair_shift) > kPairMaxDigits))
            ++r->pair_shift;
    }
    std::vector<BinView> pairviews(binviews);
    std::vector<uint32_t> xcd_order;
    for (uint32_t v = 0; v < nviews; ++v) {
        BinView& b = pairviews[v];
        b.supers_x = div_up(b.tiles_x, 1u << r->pair_shift);
        b.super_shift = r->pair_shift;
        r->pair_supers = std::max(r->pair_supers, b.supers_x * div_up(b.tiles_y, 1u << r->pair_shift));
    }
    {
        // Bands of 2^pair_shift tile rows (one supertile row), band k of
        // view v to XCD (global band index + v) % 8; each XCD's list holds
        // its bands in order, a supertile's tiles together, and the eight
        // lists interleave into the launch order (idle slots pad the short
        // ones).  With a contiguous range of tiles per XCD, the XCDs holding
        // the costlier views finished up to 1.4x later.
        std::vector<uint32_t> per[8];
        uint32_t gb = 0;
        const uint32_t side = 1u << r->pair_shift;
        for (uint32_t v = 0; v < nviews; ++v) {
            const uint32_t tx = r->bviews[v].tiles_x, ty = r->bviews[v].tiles_y;
            const uint32_t sx_n = div_up(tx, side);
            for (uint32_t y0 = 0; y0 < ty; y0 += side, ++gb) {
                const uint32_t x8 = (gb + v) % 8;
                for (uint32_t sx = 0; sx < sx_n; ++sx)
                    for (uint32_t y = y0; y < std::min(ty, y0 + side); ++y)
                        for (uint32_t x = sx * side; x < std::min(tx, (sx + 1) * side); ++x)
                            per[x8].push_back((y * tx + x) * nviews + v);
            }
        }
        size_t longest = 0;
        for (uint32_t x = 0; x < 8; ++x) longest = std::max(longest, per[x].size());
        xcd_order.assign(longest * 8, ~0u);
        for (uint32_t x = 0; x < 8; ++x)
            for (size_t q = 0; q < per[x].size(); ++q) xcd_order[q * 8 + x] = per[x][q];
    }
    if (r->tile_bits + 6 > 32 || (uint64_t)r->max_tiles * nviews >= 0xffffffffull) {
        delete r;
        return fail(c, SFK_E_ARG, "eig round: %u tiles per view too many", r->max_tiles);
    }
    const size_t vn = (size_t)nviews * n;
    auto fail_out = [&](int rc) {
        sfk_eig_round_destroy(r);
        return rc;
    };
#define RALLOC(ptr, bytes)                                                       \
    do {                                                                         \
        hipError_t _e = hipMalloc(&(ptr), (bytes));                             \
        if (_e != hipSuccess) return fail_out(hip_fail(c, _e, "eig round alloc")); \
    } while (0)
    RALLOC(r->d_views, nviews * sizeof(sfk_view));
    RALLOC(r->d_bviews, nviews * sizeof(BatchView));
    RALLOC(r->proj, vn * 16);
    RALLOC(r->proj_c, vn * 4);
    RALLOC(r->opac, (size_t)n * 4);
    RALLOC(r->keys, vn * 4);
    std::vector<uint32_t> rebase(nviews);
    bool depth_fit = true;
    {
        float bmin[3], bmax[3];
        sfk_splats_info(sp, nullptr, nullptr, bmin, bmax, nullptr, nullptr);
        for (uint32_t v = 0; v < nviews; ++v)
            depth_fit = depth_rebase(r->views[v], bmin, bmax, &rebase[v]) && depth_fit;
    }
    if (r->compact) {
        RALLOC(r->cbbox, vn * 4);
        if (depth_fit) {
            RALLOC(r->d_rebase, nviews * 4);
            r->depth_bits = 24;
        }
        RALLOC(r->d_binviews, nviews * sizeof(BinView));
        RALLOC(r->offsets, vn * 4);
        RALLOC(r->bin_sstart,
               ((size_t)std::max(r->supers, r->pair_supers * nviews) + 1) * 4);
        RALLOC(r->scan_sums, (size_t)div_up((uint64_t)r->max_tiles * nviews, kScanBlock) * 4);
        if ((uint64_t)r->max_tiles * nviews <= kTileOrderMax)
            RALLOC(r->tile_order, (size_t)r->max_tiles * nviews * 4);
        if (r->pair_supers <= kPairMaxDigits) {
            RALLOC(r->d_pairviews, nviews * sizeof(BinView));
            if (xcd_order.size() < 0xffffffffull) {
                RALLOC(r->xcd_order, xcd_order.size() * 4);
                r->xcd_order_len = (uint32_t)xcd_order.size();
            }
            RALLOC(r->pair_table, (size_t)pair_table_words(n, nviews, r->pair_supers) * 4);
            RALLOC(r->pair_table_sums,
                   (size_t)pair_table_sum_words(n, nviews, r->pair_supers) * 4);
        }
    } else {
        RALLOC(r->counts, vn * 4);
        RALLOC(r->bbox, vn * 8);
    }
    RALLOC(r->counters, (2 + (size_t)nviews * 2 * kCountSlots) * 4);
    RALLOC(r->starts, ((size_t)r->max_tiles * nviews + 1) * 4);
    RALLOC(r->resp, vn * 4);
    RALLOC(r->h0, (size_t)n * 8);
    RALLOC(r->partials, (size_t)nviews * div_up(n, kEigBlock) * 8);
    RALLOC(r->eig, nviews * 8);
    RALLOC(r->best, 4);
    RALLOC(r->emit_status, (size_t)scan_emit_status_words((uint32_t)vn) * 4);
#undef RALLOC
    if (hipHostMalloc(&r->host_out, (nviews + 1) * 8, hipHostMallocDefault) != hipSuccess)
        return fail_out(fail(c, SFK_E_HIP, "eig round: pinned output"));
    hipError_t e = hipMemcpyAsync(r->d_views, r->views.data(), nviews * sizeof(sfk_view),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_bviews, r->bviews.data(), nviews * sizeof(BatchView),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && r->d_rebase)
        e = hipMemcpyAsync(r->d_rebase, rebase.data(), nviews * 4, hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess && r->d_binviews)
        e = hipMemcpyAsync(r->d_binviews, binviews.data(), nviews * sizeof(BinView),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && r->d_pairviews)
        e = hipMemcpyAsync(r->d_pairviews, pairviews.data(), nviews * sizeof(BinView),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && r->xcd_order)
        e = hipMemcpyAsync(r->xcd_order, xcd_order.data(), xcd_order.size() * 4,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail_out(hip_fail(c, e, "eig round upload"));
    int rc = radix_scratch_create_ex(c, (uint32_t)vn, 4, r->compact ? 8 : 4, &r->sort_depth);
    if (rc) return fail_out(rc);
    sfk_radix_set_digit_bits(r->sort_depth, digit_bits);
    rc = round_alloc_isects(r, next_pow2(cap0 ? cap0 : 1));
    if (rc) return fail_out(rc);
    *out = r;
    return SFK_OK;
}

extern "C" void sfk_eig_round_destroy(sfk_eig_round* r) {
    if (!r) return;
    if (r->ctx) hipStreamSynchronize(r->ctx->stream);
    if (r->exec) hipGraphExecDestroy(r->exec);
    if (r->graph) hipGraphDestroy(r->graph);
    void* bufs[] = {r->d_views, r->d_bviews, r->proj, r->proj_c, r->opac, r->keys, r->counts, r->bbox, r->cbbox,
                    r->counters, r->starts, r->resp, r->h0, r->partials, r->eig, r->best,
                    r->emit_status, r->d_binviews, r->offsets, r->bin_sstart, r->scan_sums,
                    r->d_rebase, r->tile_order, r->pair_table, r->pair_table_sums,
                    r->d_pairviews, r->xcd_order};
    for (void* p : bufs)
        if (p) hipFree(p);
    if (r->host_out) hipHostFree(r->host_out);
    round_free_isects(r);
    sfk_radix_scratch_destroy(r->sort_depth);
    delete r;
}

// Size the isect buffers from an eager run (the count is known only on the
// device), so that the captured sequence never truncates.
static int round_fit_capacity(sfk_eig_round* r, const float* a, const float* b) {
    sfk_ctx* c = r->ctx;
    int rc = round_enqueue(r, a, b);
    if (rc) return rc;
    uint32_t total = 0;
    SFK_HIP(c, hipMemcpyAsync(&total, r->counters, 4, hipMemcpyDeviceToHost,
                              c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    if (total > r->capacity) {
        rc = round_alloc_isects(r, next_pow2(total));
        if (rc) return rc;
        if (r->exec) {
            hipGraphExecDestroy(r->exec);
            r->exec = nullptr;
        }
        if (r->graph) {
            hipGraphDestroy(r->graph);
            r->graph = nullptr;
        }
        rc = round_enqueue(r, a, b);
        if (rc) return rc;
        SFK_HIP(c, hipStreamSynchronize(c->stream));
    }
    return SFK_OK;
}

extern "C" int sfk_eig_round_set_binned(sfk_eig_round* r, int on) {
    if (!r) return SFK_E_ARG;
    if (r->binned == (on != 0)) return SFK_OK;
    r->binned = on != 0;
    // A captured round holds the other emission: capture again on next use.
    if (r->ctx) hipStreamSynchronize(r->ctx->stream);
    if (r->exec) hipGraphExecDestroy(r->exec);
    if (r->graph) hipGraphDestroy(r->graph);
    r->exec = nullptr;
    r->graph = nullptr;
    return SFK_OK;
}

extern "C" int sfk_eig_round_set_pair_raster(sfk_eig_round* r, int on) {
    if (!r) return SFK_E_ARG;
    if (r->pair_raster == (on != 0)) return SFK_OK;
    r->pair_raster = on != 0;
    if (r->ctx) hipStreamSynchronize(r->ctx->stream);
    if (r->exec) hipGraphExecDestroy(r->exec);
    if (r->graph) hipGraphDestroy(r->graph);
    r->exec = nullptr;
    r->graph = nullptr;
    return SFK_OK;
}

extern "C" int sfk_eig_round_run(sfk_ctx* c, sfk_eig_round* r, const float* a, const float* b,
                                 int use_graph, double* eig_out, uint32_t* best_out) {
    if (!c || !r || !a || !b) return SFK_E_ARG;
    int rc;
    if (!use_graph) {
        rc = round_enqueue(r, a, b);
        if (rc) return rc;
    } else {
        if (!r->exec || r->graph_a != a || r->graph_b != b) {
            if (!r->exec) {
                rc = round_fit_capacity(r, a, b);
                if (rc) return rc;
            }
            if (r->exec) hipGraphExecDestroy(r->exec);
            if (r->graph) hipGraphDestroy(r->graph);
            r->exec = nullptr;
            r->graph = nullptr;
            bool prof = c->prof;
            c->prof = false;  // no event records inside the capture
            SFK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = round_enqueue(r, a, b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            c->prof = prof;
            if (rc) {
                if (g) hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(c, e, "eig round capture");
            r->graph = g;
            SFK_HIP(c, hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0));
            r->graph_a = a;
            r->graph_b = b;
        }
        SFK_PROF(c, "eig_round_graph");
        SFK_HIP(c, hipGraphLaunch(r->exec, c->stream));
    }
    if (eig_out || best_out) {
        SFK_HIP(c, hipStreamSynchronize(c->stream));
        if (eig_out) memcpy(eig_out, r->host_out, (size_t)r->nviews * 8);
        if (best_out) memcpy(best_out, r->host_out + r->nviews, 4);
    }
    return SFK_OK;
}

extern "C" int sfk_eig_round_info(const sfk_eig_round* r, uint32_t view, const uint32_t** resp,
                                  float* scale, uint32_t* capacity, uint32_t* total_tiles) {
    if (!r || view >= r->nviews) return SFK_E_ARG;
    if (resp) *resp = r->resp + (size_t)view * r->n;
    if (scale) *scale = r->bviews[view].scale;
    if (capacity) *capacity = r->capacity;
    if (total_tiles) *total_tiles = r->total_tiles;
    return SFK_OK;
}

extern "C" int sfk_eig_round_counters(sfk_ctx* c, const sfk_eig_round* r, uint32_t* counters) {
    if (!c || !r || !counters) return SFK_E_ARG;
    const size_t words = 2 + (size_t)r->nviews * 2 * kCountSlots;
    std::vector<uint32_t> h(words);
    SFK_HIP(c, hipMemcpyAsync(h.data(), r->counters, words * 4, hipMemcpyDeviceToHost,
                              c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t v = 0; v < r->nviews; ++v) {
        uint32_t is = 0, vis = 0;
        for (uint32_t s = 0; s < kCountSlots; ++s) {
            is += h[2 + ((size_t)v * kCountSlots + s) * 2];
            vis += h[2 + ((size_t)v * kCountSlots + s) * 2 + 1];
        }
        counters[2 * v] = is;
        counters[2 * v + 1] = vis;
    }
    counters[2 * r->nviews] = h[0];
    return SFK_OK;
}

// The plain round of the reference (Segmenter::select_view, active.rs:113-129):
// per candidate, a full responsibility render then a synchronous EIG, argmax
// on the host.
extern "C" int sfk_eig_round_plain(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* scratch,
                                   const sfk_camera* cams, const uint32_t* widths,
                                   const uint32_t* heights, uint32_t nviews, const float* a,
                                   const float* b, uint32_t* resp, double* eig_out,
                                   uint32_t* best_out) {
    if (!c || !sp || !scratch || !cams || !widths || !heights || !a || !b || !resp)
        return SFK_E_ARG;
    uint32_t n = 0;
    sfk_splats_info(sp, &n, nullptr, nullptr, nullptr, nullptr, nullptr);
    uint32_t best = 0;
    double bv = -INFINITY;
    for (uint32_t v = 0; v < nviews; ++v) {
        float scale = 1.0f;
        int rc = sfk_render_responsibility(c, sp, scratch, &cams[v], widths[v], heights[v], resp,
                                           &scale);
        if (rc) return rc;
        double e = 0.0;
        rc = sfk_eig(c, a, b, resp, n, scale, &e);
        if (rc) return rc;
        if (eig_out) eig_out[v] = e;
        if (e > bv) {
            bv = e;
            best = v;
        }
    }
    if (best_out) *best_out = best;
    return SFK_OK;
}

// Autotuner: time the batched round (graph replays, HIP events) for every
// tile width in {8, 16, 32} and radix digit width in {4, 8}; report the
// fastest.  ms_out (6 entries, optional) receives the median replay time of
// each configuration in that order.
extern "C" int sfk_eig_round_autotune(sfk_ctx* c, const sfk_splats* sp, const sfk_camera* cams,
                                      const uint32_t* widths, const uint32_t* heights,
                                      uint32_t nviews, const float* a, const float* b,
                                      uint32_t reps, uint32_t* best_tile_w,
                                      uint32_t* best_digit_bits, float* ms_out) {
    if (!c || !sp || !cams || !widths || !heights || !a || !b || !best_tile_w || !best_digit_bits)
        return SFK_E_ARG;
    const uint32_t tws[3] = {8, 16, 32}, dbs[2] = {4, 8};
    float best = INFINITY;
    int k = 0;
    if (reps == 0) reps = 3;
    for (uint32_t ti = 0; ti < 3; ++ti)
        for (uint32_t di = 0; di < 2; ++di, ++k) {
            sfk_eig_round* r = nullptr;
            int rc = sfk_eig_round_create(c, sp, cams, widths, heights, nviews, tws[ti], dbs[di], &r);
            if (rc) return rc;
            rc = sfk_eig_round_run(c, r, a, b, 1, nullptr, nullptr);  // capture + warm
            std::vector<float> t;
            for (uint32_t i = 0; i < reps && !rc; ++i) {
                float ms = 0.0f;
                rc = sfk_timer_start(c);
                if (!rc) rc = sfk_eig_round_run(c, r, a, b, 1, nullptr, nullptr);
                if (!rc) rc = sfk_timer_stop_ms(c, &ms);
                t.push_back(ms);
            }
            sfk_eig_round_destroy(r);
            if (rc) return rc;
            std::sort(t.begin(), t.end());
            float med = t[t.size() / 2];
            if (ms_out) ms_out[k] = med;
            if (med < best) {
                best = med;
                *best_tile_w = tws[ti];
                *best_digit_bits = dbs[di];
            }
        }
    return SFK_OK;
}
