// Internal declarations shared by the libsfk translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sfk.h"

struct sfk_prof_rec {
    const char* name;
    hipEvent_t a, b;
};

struct sfk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // EIG reduction workspace (grown on demand).
    double* eig_partials = nullptr;
    uint32_t eig_partials_cap = 0;
    double* eig_out = nullptr;
    // Device address of the file-scope EIG partials array that
    // sfk_eig_device_noalloc reduces into (this device's copy), resolved at
    // context creation; null with the lookup's error kept when it failed.
    double* eig_fixed = nullptr;
    hipError_t eig_fixed_err = hipSuccess;
    // Per-launch event timing (sfk_prof_enable).
    bool prof = false;
    std::vector<sfk_prof_rec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
};

namespace sfk {

// Event pair around one launch when profiling is on.
hipEvent_t prof_event(sfk_ctx* c);
struct ProfScope {
    sfk_ctx* c;
    const char* name;
    hipEvent_t a = nullptr;
    ProfScope(sfk_ctx* ctx, const char* nm) : c(ctx), name(nm) {
        if (c && c->prof) {
            a = prof_event(c);
            hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope() {
        if (a) {
            hipEvent_t b = prof_event(c);
            hipEventRecord(b, c->stream);
            c->prof_recs.push_back({name, a, b});
        }
    }
};
#define SFK_PROF(ctx, name) sfk::ProfScope _sfk_prof_scope(ctx, name)

constexpr int kWave = 64;

// Set ctx->err and return code.
int fail(sfk_ctx* ctx, int code, const char* fmt, ...);
int hip_fail(sfk_ctx* ctx, hipError_t e, const char* what);

#define SFK_HIP(ctx, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return sfk::hip_fail(ctx, _e, #expr); \
    } while (0)

#define SFK_LAUNCHED(ctx, what)                                          \
    do {                                                                 \
        hipError_t _e = hipGetLastError();                               \
        if (_e != hipSuccess) return sfk::hip_fail(ctx, _e, what);       \
    } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Slot pairs of the informational (isects, visible) counts of a projection.
constexpr uint32_t kCountSlots = 64;
// Render counters: [0] total isects (written by the gather scan), [1] the
// binned emission's pair count, then kCountSlots (isects, visible) pairs
// added by the projection.
constexpr uint32_t kCounterWords = 2 + 2 * kCountSlots;

uint32_t* scan_scratch_sums(sfk_scan_scratch* s);
// Per-digit totals of the last radix pass of a single-segment argsort on `s`
// (its digit width is s's digit bits).
const uint32_t* radix_scratch_totals(sfk_radix_scratch* s, uint32_t* digit_bits);
inline uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)(p > 0xffffffffull ? 0x80000000u : p);
}

// ---- kernel launchers (each defined in the .hip file holding the kernel)

hipError_t launch_fill_u32(hipStream_t s, uint32_t* buf, uint32_t v, size_t count);
hipError_t launch_zero2_u32(hipStream_t s, uint32_t* a, size_t na, uint32_t* b, size_t nb);
hipError_t launch_copy_u32(hipStream_t s, uint32_t* dst, const uint32_t* src, size_t count);

hipError_t launch_project(hipStream_t s, const sfk_view& v, const float* attrs, const float* sh,
                          uint32_t n, uint32_t sh_per_ch, uint32_t tile_w, uint32_t* depth_keys,
                          float* proj, uint32_t* tile_counts, uint32_t* bbox, uint32_t* counters,
                          uint32_t* cbbox = nullptr, float* dcol = nullptr,
                          const float* post_a = nullptr, const float* post_b = nullptr);

// Radix sort: one pass = upsweep + digit scan + scatter.
constexpr uint32_t kSortThreads = 512;
// Smallest u64 sort (host count) that tries the bucket finish.
constexpr uint32_t kBucketMinKeys = 1u << 20;
constexpr uint64_t kSortSmallMax = 1u << 22;  // up to this many keys: 2048-key blocks
uint32_t radix_keys_per_block(uint64_t total_keys);   // 2048 or 4096
uint32_t radix_max_blocks(uint32_t max_elems);
struct SortArgs {
    const void* keys_in;
    const void* vals_in;      // nullptr: identity (index within the segment)
    const uint32_t* payload;  // 8-byte identity values: (index, payload[index])
    void* keys_out;           // nullptr: keys not written
    void* vals_out;
    uint32_t n_max;           // one segment: element count (or its cap with d_n)
    const uint32_t* d_n;      // one segment: device-side count, may be nullptr
    uint32_t nseg, seg_len;   // nseg > 1: segments of seg_len keys each
    uint32_t shift, width;
    uint32_t kpb;             // keys per block (radix_keys_per_block)
    uint32_t* hist;           // nseg * RADIX rows of `stride` blocks
    uint32_t stride;
    uint32_t* totals;         // nseg * RADIX
    // Pass 0 of u32 keys may rebase them first: key - base (base per segment
    // from rebase_bases when mode 2, rebase_base when mode 1), 0xffffffff
    // kept, the rest clamped to [0, 0xfffffe].  Order-preserving for keys in
    // [base, base + 0xfffffe]; later passes see the rebased keys.
    uint32_t rebase_mode;
    uint32_t rebase_base;
    const uint32_t* rebase_bases;
};
// Depth keys of a view (f32 bits of camera z of visible splats, 0xffffffff
// for culled ones) rebased so that 24 bits order them: see depth_rebase.
struct KeyRebase {
    uint32_t base = 0;                  // one segment
    const uint32_t* bases = nullptr;    // or one per segment (device)
};
// A lower bound `base` on the depth keys of every splat visible in view v
// whose mean lies in [bmin, bmax] (the splats' AABB); true when all of them
// lie in [base, base + 0xfffffe], i.e. a 24-bit sort of the rebased keys
// orders them as the 32-bit sort of the keys does (culled keys last).
bool depth_rebase(const sfk_view& v, const float bmin[3], const float bmax[3], uint32_t* base);
hipError_t launch_radix_pass(hipStream_t s, const SortArgs& a, uint32_t key_bytes,
                             uint32_t val_bytes, uint32_t digit_bits);
// Bucket finish of u64 sorts (sfk_sort.hip): keys grouped by their 16 bits
// at `shift` (two stable passes) -> first / last / size[65536] of every
// bucket and the largest size per group of 256 buckets
// (wg_max[bucket_size_groups()], device); then every bucket of at most
// bucket_cap() keys sorted by its bits below `shift` (lbits of them; larger
// buckets are left alone).
hipError_t launch_bucket_split(hipStream_t s, const uint64_t* grouped_keys, uint32_t n,
                               uint32_t shift, uint32_t* first, uint32_t* last, uint32_t* size,
                               uint32_t* wg_max);
uint32_t bucket_size_groups();
uint32_t bucket_cap();
hipError_t launch_bucket_sort(hipStream_t s, const uint64_t* keys_in, const uint32_t* vals_in,
                              uint64_t* keys_out, uint32_t* vals_out, const uint32_t* first,
                              const uint32_t* size, uint32_t lbits);
// Radix scratch with 8-byte values (u32 keys only), and the argsort of u32
// keys whose identity values carry a u32 payload: out[r] = (index, payload[index])
// of the element of rank r within its segment.
int radix_scratch_create_ex(sfk_ctx* c, uint32_t max_elems, uint32_t key_bytes,
                            uint32_t val_bytes, sfk_radix_scratch** out);
int argsort_payload_u32(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys,
                        const uint32_t* payload, uint32_t n, uint32_t nseg, uint32_t bits,
                        uint2** out, const KeyRebase* rebase = nullptr);
// The buffer pair of a 4-byte-key, 8-byte-value scratch that an argsort on it
// may take as input (the first pass writes the other pair), and the stable
// argsort of u32 keys carrying 8-byte values (one segment, optional device
// count; sorted keys written).
int radix_scratch_spare_u32x2(sfk_radix_scratch* s, uint32_t** keys, uint2** vals);
int argsort_u32_vals8(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys, const uint2* vals,
                      uint32_t n, const uint32_t* d_n, uint32_t bits, uint32_t** out_keys,
                      uint2** out_vals);

// Scan.
constexpr uint32_t kScanBlock = 4096;
hipError_t launch_scan(hipStream_t s, const uint32_t* gids, const uint32_t* counts,
                       uint32_t* offsets, uint32_t n, uint32_t* block_sums, uint32_t seg = 0,
                       uint32_t* total_out = nullptr);

// Fused gather scan + isect emission (scan.rs:215-255 then raster.rs:12-48).
// status_words: scan_emit_status_words(n) u32 of device scratch (cleared by
// the launch itself with a fill kernel).
struct BatchView;
uint32_t scan_emit_status_words(uint32_t n);
// Largest tile grid side whose bboxes fit the 8-bit compact form.
constexpr uint32_t kCompactGrid = 255;
// Either order + tile_counts + bbox (gathered per rank), or `ordered`:
// (splat, compact bbox) per rank straight from the payload depth sort, in
// which case order_out (optional) receives the depth order.
struct EmitSource {
    const uint32_t* order = nullptr;
    const uint32_t* tile_counts = nullptr;
    const uint32_t* bbox = nullptr;
    const uint2* ordered = nullptr;
    uint32_t* order_out = nullptr;
};
hipError_t launch_scan_emit(hipStream_t s, const EmitSource& src, uint32_t n,
                            uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                            uint32_t* gaussian_ids, uint32_t* offsets, uint32_t* total_out,
                            uint32_t* status_words, uint32_t seg_len = 0,
                            const BatchView* views = nullptr, uint32_t view_shift = 0);

// Per-view constants of a batched round (device array, one per view).
struct BatchView {
    uint32_t w, h, tiles_x, tiles_y;
    uint32_t tile_base;  // first global tile id of this view
    float scale;         // fixed-point scale 2^fractional_bits(w, h)
    uint32_t pad[2];
};

// Binned isect emission (compact grids): the isects of every tile in depth
// order, as emission + stable tile argsort + tile ranges produce them,
// without sorting the isects (sfk_bin.hip).
constexpr uint32_t kBinSuper = 8;     // tiles per supertile side
constexpr uint32_t kBinSuperShift = 3;
constexpr uint32_t kBinChunk = 512;   // sorted pairs per wavefront chunk
struct BinView {
    uint32_t tiles_x, tiles_y;  // this view's tile grid
    uint32_t supers_x;          // supertiles per row, ceil(tiles_x >> super_shift)
    uint32_t key_base;          // view << view_shift: pair keys are key_base + local supertile
    uint32_t out_mul, out_add;  // tile_starts index = local tile * out_mul + out_add
    uint32_t gid_add;           // written splat id = in-view splat + gid_add
    // log2 of the supertile side in tiles: kBinSuperShift, except for the
    // direct pair emission (sfk_bin_pairs.hip), which takes other sizes.
    uint32_t super_shift;
};
// Largest isect capacity of the binned emission.  Its prefixes saturate at
// 2^31 - 1, so a saturated total must still read as above the capacity
// (truncation on); larger capacities take emission + stable tile argsort.
constexpr uint32_t kBinMaxCap = 1u << 30;
inline uint32_t bin_supers(uint32_t tiles_x, uint32_t tiles_y) {
    return div_up(tiles_x, kBinSuper) * div_up(tiles_y, kBinSuper);
}
// Chunk slots for `cap` pairs over `supers` supertile ids: chunk k of
// supertile s is slot start[s] / kBinChunk + s + k (disjoint ranges).
inline uint32_t bin_chunk_slots(uint32_t cap, uint32_t supers) {
    return cap / kBinChunk + supers + 1;
}
// Largest tile count for which the binned emission writes a heavy-first
// tile order (the rasterizers' launch order).
constexpr uint32_t kTileOrderMax = 65536;

struct BinJob {
    const uint2* ordered;   // (in-view splat, compact bbox) per depth rank
    uint32_t n;             // ranks (all views)
    uint32_t seg_len;       // ranks per view when batched, else 0
    const BinView* views;   // device, one per view (batched), or nullptr:
    BinView view0;          //   the single view, by value
    uint32_t nviews;
    uint32_t view_shift;    // bits of the local supertile ids (the sorted bits), >= 1
    uint32_t supers;        // supertile ids: local * nviews + view, max local x nviews
    uint32_t out_tiles;     // tile_starts entries before the total
    bool zero_starts;       // some tile_starts entries belong to no supertile
    uint32_t cap;           // isect capacity (the reference's max_isects), 2..kBinMaxCap
    uint32_t* order_out;    // optional: depth order
    uint32_t* offsets;      // optional, per rank: exclusive isect offset (n words)
    uint32_t* offs_by_splat;  // n words: the isect offset of every rank whose
                              // isects may be truncated, at its (view-major) splat
    sfk_radix_scratch* sort;  // >= cap elements, 4-byte keys, 8-byte values; the
                              // (supertile, splat | bbox) pairs live in its spare pair
    uint32_t* sstart;       // supers + 1
    uint32_t* chunk_map;    // 4 words per chunk slot: bin_chunk_slots(cap, supers)
    uint32_t* chunk_counts; // 64 per chunk slot
    uint32_t* scan_sums;    // div_up(out_tiles, kScanBlock) words
    uint32_t* tile_starts;  // out_tiles + 1
    uint32_t* tile_order;   // optional, out_tiles: tiles heaviest first (written only
                            // when out_tiles <= kTileOrderMax)
    uint32_t* gids;         // cap: splat per isect
    uint32_t* tile_ids;     // cap or nullptr: tile per isect
    uint32_t* counters;     // [0] raw isect count, [1] pair count
    uint32_t* status;       // scan_emit_status_words(n)
    bool status_zeroed;     // the caller cleared `status` (else bin_emit does)
    bool self_clean;        // leave `status` clear once the emission scan is done with
                            // it (the next job then needs no clearing), and zero
                            // clear[0, clear_n) with it
    uint32_t* clear;
    uint32_t clear_n;
    uint32_t* host_total;   // optional device view of mapped pinned host memory:
                            // receives the raw isect count as well
    // Pairs only: stop after the pair sort and the supertile ranges (sstart);
    // *pairs_out receives the sorted pair values.  chunk_map, chunk_counts,
    // scan_sums, tile_starts, tile_order, gids and tile_ids are unused.
    bool pairs_only;
    const uint2** pairs_out;
    // Pairs only, optional: with at most kPairMaxDigits local supertiles per
    // view, the pairs are written straight into supertile order (no look-back,
    // pair sort or ranges pass; sfk_bin_pairs.hip) using these
    // pair_table_words / pair_table_sum_words words of scratch.
    uint32_t* pair_table;
    uint32_t* pair_table_sums;
    // Skip the placement pass: counts, tile starts and the tile order as
    // usual, *pairs_out the sorted pairs; the rasterizers place what they
    // consume (RasterArgs::gids_out).
    bool skip_place;
    // Skip place with a tile order: 65 words of class counts (zero before the
    // first prepare; each emission leaves them zero) and one word per tile.
    uint32_t* class_count;
    uint32_t* tile_cls;
};
int bin_emit(sfk_ctx* c, const BinJob& j);
constexpr uint32_t kPairMaxDigits = 256;
// Smallest supertile side, in pixels, of the batched round's direct pair
// emission.
#ifndef SFK_PAIR_SUPER_PX
#define SFK_PAIR_SUPER_PX 64
#endif
constexpr uint32_t kPairSuperPx = SFK_PAIR_SUPER_PX;
// Most supertiles a row of the direct pair emission's views (its pair write
// ranks a bbox's supertiles by a reciprocal table of the row spans).
constexpr uint32_t kPairMaxSpan = 32;
uint32_t pair_table_words(uint32_t n_per_view, uint32_t views, uint32_t digits);
uint32_t pair_table_sum_words(uint32_t n_per_view, uint32_t views, uint32_t digits);
int bin_pairs_direct(sfk_ctx* c, const BinJob& j, uint32_t digits);

// Raster stage.
hipError_t launch_map_isects(hipStream_t s, const uint32_t* order, const uint32_t* tile_counts,
                             const uint32_t* bbox, const uint32_t* offsets, uint32_t n,
                             uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                             uint32_t* gaussian_ids, uint32_t seg_len = 0,
                             const BatchView* views = nullptr);
// nviews > 0: keys are view << view_shift | local tile, sorted by local tile
// (then view, from the view-major input); starts are indexed by
// local * nviews + view over num_tiles = max local tiles * nviews.
hipError_t launch_tile_starts(hipStream_t s, const uint32_t* tile_ids, uint32_t n,
                              const uint32_t* d_n, uint32_t num_tiles, uint32_t* starts,
                              uint32_t nviews = 0, uint32_t view_shift = 0);

enum RasterMode { kRgb = 0, kTotal = 1, kEvidence = 2 };
struct RasterArgs {
    uint32_t w, h, tiles_x, tiles_y, tile_w, row_stride;
    float scale;
    const uint32_t* tile_starts;
    const uint32_t* tile_order;  // optional launch order of the tiles (heaviest first)
    const uint32_t* gids;
    const float* proj;
    const uint32_t* mask;
    uint32_t words_per_row;
    float* rgba;
    uint32_t* bitmap;
    float* final_t;
    uint32_t* n_contrib;
    uint32_t* resp;
    uint32_t* fg;
    uint32_t* bg;
    // batched responsibility (launch_rasterize_batch): proj holds one float4
    // per (view, splat) (mean x, y, conic a, b), proj_c conic c per (view,
    // splat) and opac the opacity per splat.
    const BatchView* views;
    uint32_t nviews, batch_tiles, seg_len;
    const float* proj_c;
    const float* opac;
    // Batched, pair source (optional): no isect lists; each tile takes its
    // isects straight from its supertile's sorted (splat | may-truncate << 31,
    // compact bbox) pairs, pairs[pair_starts[s * nviews + view] ..
    // pair_starts[s * nviews + view + 1]) for local supertile s, keeping the
    // pairs whose bbox covers the tile (and, for a flagged pair, whose
    // emission slot offs_by_splat[view-major splat] + index in bbox is below
    // cap).
    const uint2* pairs;
    const uint32_t* pair_starts;
    const uint32_t* offs_by_splat;
    uint32_t cap;
    // Batched pair source: log2 of the supertile side in tiles (the pair
    // emission's BinView::super_shift).
    uint32_t super_shift;
    // One view, pair source: the rasterizers write the isects they consume
    // to gids_out at their slots of the tile ranges (tile_starts), so that a
    // backward pass finds every isect up to each pixel's last contributor.
    uint32_t* gids_out;
};
hipError_t launch_rasterize(hipStream_t s, int mode, const RasterArgs& a);
hipError_t launch_rasterize_batch(hipStream_t s, const RasterArgs& a);

// Backward.
uint32_t loss_num_partials(uint32_t np);
hipError_t launch_loss_l2(hipStream_t s, const float* img, const float* target, uint32_t np,
                          float* d_img, double* partials, double* out);
// With d_img == nullptr the L2 loss is fused: dL/dimg = 2 (img - target)
// and one f64 partial per tile in the adjoint raster; the projection adjoint
// (launched next with the same BwdLoss) writes their fixed-order total.
struct BwdLoss {
    const float* img;
    const float* target;
    double* partials;  // one per tile
    uint32_t num_partials;
    double* out;
};
hipError_t launch_raster_bwd(hipStream_t s, const RasterArgs& ra, const float* final_t,
                             const uint32_t* n_contrib, const float* d_img, uint32_t n,
                             float* grad2d, const BwdLoss* loss = nullptr);
hipError_t launch_project_bwd(hipStream_t s, const sfk_view& v, const float* attrs,
                              const float* sh, uint32_t n, uint32_t sh_per_ch,
                              float* grad2d, float* g_attrs, float* g_sh,
                              const BwdLoss* loss = nullptr, const float* dcol = nullptr);

// Seg.
hipError_t launch_eig(hipStream_t s, const float* a, const float* b, const uint32_t* resp,
                      uint32_t n, float scale, double* partials, double* out);
uint32_t eig_num_partials(uint32_t n);
hipError_t launch_beta_update(hipStream_t s, float* a, float* b, const uint32_t* fg,
                              const uint32_t* bg, uint32_t n, float scale);
// paint.rs kernels: DC tint of the a > b splats, f32 copy (the posterior
// paint is fused into the projection, launch_project's post_a / post_b).
hipError_t launch_tint_colors(hipStream_t s, float* sh, const float* a, const float* b,
                              uint32_t n, float r, float g, float bl);
hipError_t launch_copy_f32(hipStream_t s, float* dst, const float* src, size_t count);
// PLY vertex rows (l.stride floats each, as in the file) -> field-major planes.
hipError_t launch_ply_decode(hipStream_t s, const sfk_ply_layout& l, const float* rows,
                             float* attrs, float* sh);
// views::localize on the device; partials: 4 * localize_blocks(n) doubles;
// out (device): [cx, cy, cz, sum of weights, radius].
uint32_t localize_blocks(uint32_t n);
hipError_t launch_localize(hipStream_t s, const float* attrs, const float* a, const float* b,
                           uint32_t n, double* partials, double* out);

}  // namespace sfk
