// Render and seg orchestration behind the C ABI: Splats, RenderScratch,
// prepare_isects / finalize / render_with (src/render.rs:114-460) and the
// responsibility, EIG and Beta-update entry points (src/seg/mod.rs:146-210,
// src/seg/beta.rs:240-286).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "sfk_internal.h"

using namespace sfk;

namespace {

// render.rs:12-18
constexpr uint32_t kIsectsUpperBound = 2u * 512u * 65535u;
constexpr uint32_t kInitialIsectsCap = 1u << 22;

}  // namespace

struct sfk_splats {
    sfk_ctx* ctx = nullptr;
    uint32_t n = 0, sh_per_ch = 0;
    float* attrs = nullptr;
    float* sh = nullptr;
    float bmin[3], bmax[3];
};

extern "C" int sfk_splats_create(sfk_ctx* c, const float* attrs, const float* sh, uint32_t n,
                                 uint32_t sh_per_ch, sfk_splats** out) {
    if (!c || !out) return SFK_E_ARG;
    *out = nullptr;
    if (n == 0) return fail(c, SFK_E_ARG, "Splats::new: zero splats");
    if (!attrs || !sh || !(sh_per_ch == 1 || sh_per_ch == 4 || sh_per_ch == 9 || sh_per_ch == 16))
        return fail(c, SFK_E_ARG, "splats: bad attributes/sh (sh_per_ch=%u)", sh_per_ch);
    auto* s = new (std::nothrow) sfk_splats();
    if (!s) return SFK_E_NOMEM;
    s->ctx = c;
    s->n = n;
    s->sh_per_ch = sh_per_ch;
    size_t ab = (size_t)n * SFK_ATTR_PLANES * 4, sb = (size_t)n * sh_per_ch * 3 * 4;
    hipError_t e = hipMalloc(&s->attrs, ab);
    if (e == hipSuccess) e = hipMalloc(&s->sh, sb);
    if (e == hipSuccess) e = hipMemcpyAsync(s->attrs, attrs, ab, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->sh, sh, sb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (s->attrs) hipFree(s->attrs);
        if (s->sh) hipFree(s->sh);
        delete s;
        return hip_fail(c, e, "splats upload");
    }
    // Scene AABB of the xyz planes (render.rs:253-259).
    for (int k = 0; k < 3; ++k) {
        s->bmin[k] = 3.40282347e38f;
        s->bmax[k] = -3.40282347e38f;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            float v = attrs[(size_t)k * n + i];
            s->bmin[k] = std::min(s->bmin[k], v);
            s->bmax[k] = std::max(s->bmax[k], v);
        }
    *out = s;
    return SFK_OK;
}

extern "C" int sfk_splats_create_device(sfk_ctx* c, const float* d_attrs, const float* d_sh,
                                        uint32_t n, uint32_t sh_per_ch, sfk_splats** out) {
    if (!c || !out) return SFK_E_ARG;
    *out = nullptr;
    if (n == 0) return fail(c, SFK_E_ARG, "Splats::new: zero splats");
    if (!d_attrs || !d_sh || !(sh_per_ch == 1 || sh_per_ch == 4 || sh_per_ch == 9 || sh_per_ch == 16))
        return fail(c, SFK_E_ARG, "splats: bad attributes/sh (sh_per_ch=%u)", sh_per_ch);
    auto* s = new (std::nothrow) sfk_splats();
    if (!s) return SFK_E_NOMEM;
    s->ctx = c;
    s->n = n;
    s->sh_per_ch = sh_per_ch;
    const size_t ab = (size_t)n * SFK_ATTR_PLANES * 4, sb = (size_t)n * sh_per_ch * 3 * 4;
    std::vector<float> xyz((size_t)3 * n);
    hipError_t e = hipMalloc(&s->attrs, ab);
    if (e == hipSuccess) e = hipMalloc(&s->sh, sb);
    if (e == hipSuccess) e = hipMemcpyAsync(s->attrs, d_attrs, ab, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->sh, d_sh, sb, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(xyz.data(), d_attrs, xyz.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (s->attrs) hipFree(s->attrs);
        if (s->sh) hipFree(s->sh);
        delete s;
        return hip_fail(c, e, "splats copy");
    }
    for (int k = 0; k < 3; ++k) {
        s->bmin[k] = 3.40282347e38f;
        s->bmax[k] = -3.40282347e38f;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const float v = xyz[(size_t)k * n + i];
            s->bmin[k] = std::min(s->bmin[k], v);
            s->bmax[k] = std::max(s->bmax[k], v);
        }
    *out = s;
    return SFK_OK;
}

extern "C" int sfk_splats_from_ply(sfk_ctx* c, const void* bytes, size_t len,
                                   sfk_splats** out) {
    if (!c || !bytes || !out) return SFK_E_ARG;
    *out = nullptr;
    sfk_ply_layout l;
    char err[256];
    int rc = sfk_ply_parse_header(bytes, len, &l, err, (int)sizeof err);
    if (rc) return fail(c, rc, "%s", err);
    if (!(l.sh_planes == 3 || l.sh_planes == 12 || l.sh_planes == 27 || l.sh_planes == 48))
        return fail(c, SFK_E_ARG, "PLY with %u f_rest properties: SH degree 0-3 only", l.n_rest);
    if (l.n > 0xffffffffull) return fail(c, SFK_E_ARG, "PLY with %llu vertices",
                                         (unsigned long long)l.n);
    const uint32_t n = (uint32_t)l.n;
    const size_t row_bytes = (size_t)l.stride * 4 * n;
    auto* s = new (std::nothrow) sfk_splats();
    if (!s) return SFK_E_NOMEM;
    s->ctx = c;
    s->n = n;
    s->sh_per_ch = l.sh_planes / 3;
    float* rows = nullptr;
    const char* data = static_cast<const char*>(bytes) + l.data_offset;
    hipError_t e = hipMalloc(&s->attrs, (size_t)n * SFK_ATTR_PLANES * 4);
    if (e == hipSuccess) e = hipMalloc(&s->sh, (size_t)n * l.sh_planes * 4);
    if (e == hipSuccess) e = hipMalloc(&rows, row_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(rows, data, row_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_ply_decode(c->stream, l, rows, s->attrs, s->sh);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (rows) hipFree(rows);
    if (e != hipSuccess) {
        if (s->attrs) hipFree(s->attrs);
        if (s->sh) hipFree(s->sh);
        delete s;
        return hip_fail(c, e, "PLY upload");
    }
    // Scene AABB (render.rs:253-259) straight from the rows' x, y, z.
    for (int k = 0; k < 3; ++k) {
        s->bmin[k] = 3.40282347e38f;
        s->bmax[k] = -3.40282347e38f;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            float v;
            memcpy(&v, data + ((size_t)i * l.stride + l.base[k]) * 4, 4);
            s->bmin[k] = std::min(s->bmin[k], v);
            s->bmax[k] = std::max(s->bmax[k], v);
        }
    *out = s;
    return SFK_OK;
}

extern "C" void sfk_splats_destroy(sfk_splats* s) {
    if (!s) return;
    if (s->attrs) hipFree(s->attrs);
    if (s->sh) hipFree(s->sh);
    delete s;
}

extern "C" int sfk_splats_info(const sfk_splats* s, uint32_t* n, uint32_t* sh_per_ch,
                               float bmin[3], float bmax[3], const float** d_attrs,
                               const float** d_sh) {
    if (!s) return SFK_E_ARG;
    if (n) *n = s->n;
    if (sh_per_ch) *sh_per_ch = s->sh_per_ch;
    if (bmin) memcpy(bmin, s->bmin, 12);
    if (bmax) memcpy(bmax, s->bmax, 12);
    if (d_attrs) *d_attrs = s->attrs;
    if (d_sh) *d_sh = s->sh;
    return SFK_OK;
}

struct sfk_render_scratch {
    sfk_ctx* ctx = nullptr;
    uint32_t n = 0, w = 0, h = 0, tile_w = 16, tiles_x = 0, tiles_y = 0, row_stride = 0;
    uint32_t capacity = 0, digit_bits = 8;
    uint32_t* depth_keys = nullptr;
    const uint32_t* depth_order = nullptr;  // aliases sort_depth
    float* projected = nullptr;
    uint32_t* tile_counts = nullptr;
    uint32_t* bbox = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* counters = nullptr;  // [isects_raw, visible, pad, pad]
    uint32_t* tile_ids = nullptr;
    uint32_t* gaussian_ids = nullptr;
    uint32_t* tile_starts = nullptr;
    // Tiles heaviest first, written by the binned emission of grids of at
    // most kTileOrderMax tiles; order_valid while the last prepare wrote it.
    uint32_t* tile_order = nullptr;
    bool order_valid = false;
    uint32_t* bitmap = nullptr;
    float* rgba = nullptr;
    float* final_t = nullptr;
    uint32_t* n_contrib = nullptr;
    sfk_radix_scratch* sort_depth = nullptr;
    sfk_radix_scratch* sort_tile = nullptr;
    uint32_t* emit_status = nullptr;  // look-back words of the fused scan/emission
    // Two banks of kCounterWords: a prepare counts into bank `bank`, and
    // the binned emission's scan clears the other bank's count slots and
    // its own look-back words for the next prepare (emit_clean), so a
    // binned frame starts with no clearing launch.
    uint32_t bank = 0;
    uint32_t* cur_counters = nullptr;  // the bank of the last prepare
    bool emit_clean = false;
    uint32_t* cbbox = nullptr;        // compact tile bbox per splat (depth-sort payload)
    uint32_t* order_buf = nullptr;    // depth order written by the emission
    // Previous frame's raw emission, read back asynchronously for growth.
    uint32_t* host_counters = nullptr;
    uint32_t* host_counters_dev = nullptr;  // device view of the mapped pinned word
    hipEvent_t counters_ready = nullptr;
    bool counters_pending = false;
    uint32_t last_isects_raw = 0;
    // Current frame outputs (sorted isects alias sort_tile).
    const uint32_t* sorted_tiles = nullptr;
    const uint32_t* sorted_gids = nullptr;
    sfk_view view;  // view of the last prepare
    bool has_view = false;
    bool has_colors = false;  // the last prepare wrote colours (SH or posterior)
    bool sh_colors = false;   // ... and they are this frame's SH colours
    float* grad2d = nullptr;  // 9 planes x n screen-space gradients (backward)
    bool grad2d_dirty = true; // grad2d not known to be all zero
    double* loss_partials = nullptr;  // one per tile (fused L2 backward)
    uint32_t loss_partials_cap = 0;
    // Binned emission (compact grids): supertile ranges, chunk map, per-chunk
    // tile counts (the last two sized by the isect capacity).
    bool binned = true;
    bool want_tile_ids = true;  // the binned emission also writes tile_ids
    // Pair raster: the binned emission stops before placing the isects; the
    // rasterizers take each tile's isects from its supertile's sorted pairs
    // and write back the ones they consume (pairs_ready while the last
    // prepare left its pairs for them).
    bool pair_raster = false;
    bool pairs_ready = false;
    const uint2* pairs = nullptr;
    uint32_t* bin_sstart = nullptr;
    uint32_t* bin_map = nullptr;
    uint32_t* bin_counts = nullptr;
    uint32_t* bin_offs = nullptr;  // isect offsets of truncation candidates, by splat
    // Pair raster: work-class counts (65 words, kept zero between prepares)
    // and per-tile class + index, from which the tile order is written.
    uint32_t* class_count = nullptr;
    uint32_t* tile_cls = nullptr;
    // d colour / d view direction per splat (9 planes), kept by the RGB
    // projection once a backward has run, so the next adjoint reads 9 floats
    // per splat instead of the SH coefficients.
    float* dcol = nullptr;
    bool want_dcol = false;
    bool has_dcol = false;  // the last prepare wrote dcol
};

static bool compact_grid(const sfk_render_scratch* s) {
    return s->tiles_x <= kCompactGrid && s->tiles_y <= kCompactGrid;
}

static void scratch_free_frame(sfk_render_scratch* s) {
    void* bufs[] = {s->depth_keys, s->projected, s->tile_counts, s->bbox, s->offsets,
                    s->counters, s->tile_starts, s->bitmap, s->rgba, s->final_t, s->n_contrib,
                    s->tile_ids, s->gaussian_ids, s->grad2d, s->emit_status, s->cbbox,
                    s->order_buf, s->bin_sstart, s->bin_map, s->bin_counts, s->bin_offs,
                    s->dcol, s->tile_order, s->class_count, s->tile_cls};
    for (void* p : bufs)
        if (p) hipFree(p);
    s->depth_keys = nullptr; s->projected = nullptr; s->tile_counts = nullptr; s->bbox = nullptr;
    s->offsets = nullptr; s->counters = nullptr; s->tile_starts = nullptr; s->bitmap = nullptr;
    s->rgba = nullptr; s->final_t = nullptr; s->n_contrib = nullptr; s->tile_ids = nullptr;
    s->gaussian_ids = nullptr;
    s->tile_order = nullptr;
    s->order_valid = false;
    s->pairs_ready = false;
    s->grad2d = nullptr;
    s->grad2d_dirty = true;
    s->emit_status = nullptr;
    s->cbbox = s->order_buf = nullptr;
    s->bin_sstart = s->bin_map = s->bin_counts = s->bin_offs = nullptr;
    s->class_count = s->tile_cls = nullptr;
    s->dcol = nullptr;
    s->has_dcol = false;
    s->has_view = false;
    s->has_colors = false;
    s->sh_colors = false;
    sfk_radix_scratch_destroy(s->sort_depth);
    sfk_radix_scratch_destroy(s->sort_tile);
    s->sort_depth = s->sort_tile = nullptr;
    s->depth_order = s->sorted_tiles = s->sorted_gids = nullptr;
}

static int alloc_isects(sfk_render_scratch* s, uint32_t cap) {
    sfk_ctx* c = s->ctx;
    void* bufs[] = {s->tile_ids, s->gaussian_ids, s->bin_map, s->bin_counts};
    for (void* p : bufs)
        if (p) hipFree(p);
    s->tile_ids = s->gaussian_ids = s->bin_map = s->bin_counts = nullptr;
    sfk_radix_scratch_destroy(s->sort_tile);
    s->sort_tile = nullptr;
    s->sorted_tiles = s->sorted_gids = nullptr;
    SFK_HIP(c, hipMalloc(&s->tile_ids, (size_t)cap * 4));
    SFK_HIP(c, hipMalloc(&s->gaussian_ids, (size_t)cap * 4));
    if (compact_grid(s) && cap <= kBinMaxCap) {
        const size_t slots = bin_chunk_slots(cap, bin_supers(s->tiles_x, s->tiles_y));
        SFK_HIP(c, hipMalloc(&s->bin_map, slots * 16));
        SFK_HIP(c, hipMalloc(&s->bin_counts, slots * 64 * 4));
    }
    // 8-byte values where the binned emission may sort its pairs.
    const bool bin = compact_grid(s) && cap <= kBinMaxCap;
    int rc = radix_scratch_create_ex(c, cap, 4, bin ? 8 : 4, &s->sort_tile);
    if (rc) return rc;
    sfk_radix_set_digit_bits(s->sort_tile, s->digit_bits);
    s->capacity = cap;
    return SFK_OK;
}

// RenderScratch::new (render.rs:143-167).
static int scratch_build(sfk_render_scratch* s, uint32_t n, uint32_t w, uint32_t h) {
    sfk_ctx* c = s->ctx;
    scratch_free_frame(s);
    s->n = n; s->w = w; s->h = h;
    s->tiles_x = (w + s->tile_w - 1) / s->tile_w;
    s->tiles_y = (h + s->tile_w - 1) / s->tile_w;
    s->row_stride = ((w * 4 + 255) / 256 * 256) / 4;
    uint64_t tiles = (uint64_t)s->tiles_x * s->tiles_y;
    size_t nn = n ? n : 1, px = (size_t)w * h;
    SFK_HIP(c, hipMalloc(&s->depth_keys, nn * 4));
    SFK_HIP(c, hipMalloc(&s->projected, nn * SFK_PROJ_STRIDE * 4));
    SFK_HIP(c, hipMalloc(&s->tile_counts, nn * 4));
    SFK_HIP(c, hipMalloc(&s->bbox, nn * 8));
    SFK_HIP(c, hipMalloc(&s->offsets, nn * 4));
    SFK_HIP(c, hipMalloc(&s->counters, 2 * kCounterWords * 4));
    s->cur_counters = s->counters;
    s->bank = 0;
    s->emit_clean = false;
    SFK_HIP(c, hipMalloc(&s->tile_starts, (size_t)(tiles + 1) * 4));
    if (tiles <= kTileOrderMax) SFK_HIP(c, hipMalloc(&s->tile_order, (size_t)tiles * 4));
    SFK_HIP(c, hipMalloc(&s->bitmap, (size_t)h * s->row_stride * 4));
    SFK_HIP(c, hipMalloc(&s->rgba, px * 16));
    SFK_HIP(c, hipMalloc(&s->final_t, px * 4));
    SFK_HIP(c, hipMalloc(&s->n_contrib, px * 4));
    SFK_HIP(c, hipMalloc(&s->cbbox, nn * 4));
    SFK_HIP(c, hipMalloc(&s->order_buf, nn * 4));
    int rc = radix_scratch_create_ex(c, n, 4, 8, &s->sort_depth);
    if (rc) return rc;
    sfk_radix_set_digit_bits(s->sort_depth, s->digit_bits);
    SFK_HIP(c, hipMalloc(&s->emit_status, (size_t)scan_emit_status_words(n) * 4));
    if (compact_grid(s)) {
        SFK_HIP(c, hipMalloc(&s->bin_sstart,
                             ((size_t)bin_supers(s->tiles_x, s->tiles_y) + 1) * 4));
        SFK_HIP(c, hipMalloc(&s->bin_offs, nn * 4));
        SFK_HIP(c, hipMalloc(&s->class_count, 65 * 4));
        SFK_HIP(c, hipMemsetAsync(s->class_count, 0, 65 * 4, c->stream));
        SFK_HIP(c, hipMalloc(&s->tile_cls, (size_t)tiles * 4));
    }
    uint64_t cap = std::min<uint64_t>(tiles * n, kInitialIsectsCap);
    return alloc_isects(s, next_pow2(cap ? cap : 1));
}

extern "C" int sfk_render_scratch_create(sfk_ctx* c, uint32_t n, uint32_t w, uint32_t h,
                                         sfk_render_scratch** out) {
    if (!c || !out) return SFK_E_ARG;
    *out = nullptr;
    if (n == 0 || w == 0 || h == 0) return fail(c, SFK_E_ARG, "render scratch: empty shape");
    auto* s = new (std::nothrow) sfk_render_scratch();
    if (!s) return SFK_E_NOMEM;
    s->ctx = c;
    int rc = SFK_OK;
    if (hipHostMalloc(&s->host_counters, 16, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&s->host_counters_dev), s->host_counters,
                                0) != hipSuccess ||
        hipEventCreateWithFlags(&s->counters_ready, hipEventDisableTiming) != hipSuccess)
        rc = fail(c, SFK_E_HIP, "render scratch: pinned counters");
    if (!rc) rc = scratch_build(s, n, w, h);
    if (rc) {
        sfk_render_scratch_destroy(s);
        return rc;
    }
    *out = s;
    return SFK_OK;
}

extern "C" void sfk_render_scratch_destroy(sfk_render_scratch* s) {
    if (!s) return;
    if (s->ctx) hipStreamSynchronize(s->ctx->stream);
    scratch_free_frame(s);
    if (s->host_counters) hipHostFree(s->host_counters);
    if (s->loss_partials) hipFree(s->loss_partials);
    if (s->counters_ready) hipEventDestroy(s->counters_ready);
    delete s;
}

extern "C" int sfk_render_scratch_set_tile_width(sfk_render_scratch* s, uint32_t tw) {
    if (!s || !(tw == 8 || tw == 16 || tw == 32)) return SFK_E_ARG;
    if (tw == s->tile_w) return SFK_OK;
    s->tile_w = tw;
    uint32_t raw = s->last_isects_raw;
    int rc = scratch_build(s, s->n, s->w, s->h);
    s->last_isects_raw = raw;
    return rc;
}

extern "C" int sfk_render_scratch_set_binned(sfk_render_scratch* s, int on) {
    if (!s) return SFK_E_ARG;
    s->binned = on != 0;
    return SFK_OK;
}

extern "C" int sfk_render_scratch_set_tile_ids(sfk_render_scratch* s, int on) {
    if (!s) return SFK_E_ARG;
    s->want_tile_ids = on != 0;
    return SFK_OK;
}

extern "C" int sfk_render_scratch_set_pair_raster(sfk_render_scratch* s, int on) {
    if (!s) return SFK_E_ARG;
    s->pair_raster = on != 0;
    return SFK_OK;
}

extern "C" int sfk_render_scratch_set_digit_bits(sfk_render_scratch* s, uint32_t bits) {
    if (!s || bits < 4 || bits > 8) return SFK_E_ARG;
    s->digit_bits = bits;
    if (s->sort_depth) sfk_radix_set_digit_bits(s->sort_depth, bits);
    if (s->sort_tile) sfk_radix_set_digit_bits(s->sort_tile, bits);
    return SFK_OK;
}

static int grow_to(sfk_render_scratch* s, uint64_t want) {
    uint32_t cap = std::min<uint64_t>(next_pow2(want ? want : 1), kIsectsUpperBound);
    if (cap > s->capacity) return alloc_isects(s, cap);
    return SFK_OK;
}

extern "C" int sfk_render_scratch_reserve_isects(sfk_render_scratch* s, uint32_t isects) {
    if (!s) return SFK_E_ARG;
    return grow_to(s, isects);
}

extern "C" int sfk_render_scratch_buffers(const sfk_render_scratch* s, sfk_render_buffers* o) {
    if (!s || !o) return SFK_E_ARG;
    o->depth_keys = s->depth_keys;
    o->depth_order = s->depth_order;
    o->projected = s->projected;
    o->tile_counts = s->tile_counts;
    o->bbox = s->bbox;
    o->offsets = s->offsets;
    o->counters = s->cur_counters;
    o->bitmap = s->bitmap;
    o->rgba = s->rgba;
    o->final_t = s->final_t;
    o->n_contrib = s->n_contrib;
    o->n = s->n; o->w = s->w; o->h = s->h; o->row_stride = s->row_stride;
    o->capacity = s->capacity;
    return SFK_OK;
}

// RenderScratch::prepare (render.rs:169-197): rebuild on a shape change, then
// grow the isect buffers from the previous frame's raw emission.
static int scratch_prepare(sfk_render_scratch* s, uint32_t n, uint32_t w, uint32_t h) {
    sfk_ctx* c = s->ctx;
    if (s->counters_pending) {
        SFK_HIP(c, hipEventSynchronize(s->counters_ready));
        s->last_isects_raw = s->host_counters[0];
        s->counters_pending = false;
    }
    uint32_t raw = s->last_isects_raw;
    if (n != s->n || w != s->w || h != s->h) {
        int rc = scratch_build(s, n, w, h);
        if (rc) return rc;
    }
    return grow_to(s, raw);
}

static void fill_isects(const sfk_render_scratch* s, sfk_tile_isects* o) {
    o->tile_ids = s->sorted_tiles;
    o->gaussian_ids = s->sorted_gids;
    o->tile_starts = s->tile_starts;
    o->d_counters = s->cur_counters;
    o->capacity = s->capacity;
    o->tiles_x = s->tiles_x;
    o->tiles_y = s->tiles_y;
    o->tile_w = s->tile_w;
    o->row_stride = s->row_stride;
    o->w = s->w;
    o->h = s->h;
}

// Splats::prepare_isects (render.rs:271-364) without the host readback: the
// depth sort runs over all n splats (culled ones carry key 0xffffffff and
// sort last, contributing no tiles), the gather scan over all n ranks, and
// the tile sort over the device-side isect count.
static int prepare_impl(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                        const sfk_view* view, uint32_t w, uint32_t h, sfk_tile_isects* out,
                        bool colors, const float* post_a = nullptr,
                        const float* post_b = nullptr);

extern "C" int sfk_prepare_isects_view(sfk_ctx* c, const sfk_splats* sp,
                                       sfk_render_scratch* s, const sfk_view* view, uint32_t w,
                                       uint32_t h, sfk_tile_isects* out) {
    return prepare_impl(c, sp, s, view, w, h, out, true);
}

// colors = false projects geometry only (no SH): the responsibility modes
// never read the colour, and skipping it leaves every epsilon word unchanged.
// post_a / post_b (colors false): colour rows hold the posterior mean.
static int prepare_impl(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                        const sfk_view* view, uint32_t w, uint32_t h, sfk_tile_isects* out,
                        bool colors, const float* post_a, const float* post_b) {
    if (!c || !sp || !s || !view || !out) return SFK_E_ARG;
    if ((uint32_t)view->img[0] != w || (uint32_t)view->img[1] != h)
        return fail(c, SFK_E_ARG, "view is for %gx%g, not %ux%u", view->img[0], view->img[1], w, h);
    int rc = scratch_prepare(s, sp->n, w, h);
    if (rc) return rc;
    if ((uint32_t)view->tile_bounds[0] != s->tiles_x || (uint32_t)view->tile_bounds[1] != s->tiles_y)
        return fail(c, SFK_E_ARG, "view tile bounds do not match the scratch tile width %u", s->tile_w);
    const uint32_t n = sp->n;
    const uint32_t tiles = s->tiles_x * s->tiles_y;
    const bool compact = compact_grid(s);
    s->order_valid = false;
    s->pairs_ready = false;

    // Counters (and, for the binned emission, its look-back words) start at
    // 0: cleared here, or left clear by the last binned prepare.
    const bool binned = compact && s->binned && s->capacity >= 2 && s->capacity <= kBinMaxCap;
    const uint32_t bank = s->bank;
    uint32_t* ctr = s->counters + (size_t)bank * kCounterWords;
    s->cur_counters = ctr;
    if (!(binned && s->emit_clean)) {
        SFK_HIP(c, launch_zero2_u32(c->stream, s->counters, 2 * kCounterWords, s->emit_status,
                                    binned ? scan_emit_status_words(n) : 0u));
    }
    s->emit_clean = false;
    {
        SFK_PROF(c, "project");
        float* dcol = colors && s->want_dcol && sp->sh_per_ch > 1 ? s->dcol : nullptr;
        SFK_HIP(c, launch_project(c->stream, *view, sp->attrs, sp->sh, n,
                                  colors ? sp->sh_per_ch : 0u, s->tile_w, s->depth_keys,
                                  s->projected, s->tile_counts, s->bbox, ctr + 2,
                                  compact ? s->cbbox : nullptr, dcol, post_a, post_b));
        s->has_colors = colors || post_a;
        s->sh_colors = colors && !post_a;
        s->has_dcol = dcol != nullptr;
    }

    // Depth order.  On grids of at most 255 x 255 tiles the compact bbox
    // rides through the sort, and the emission reads (splat, bbox) in depth
    // order instead of gathering counts and boxes per rank.
    EmitSource src;
    if (compact) {
        // Where the splats' AABB bounds this view's visible depths to a 2^24
        // span of f32 bit patterns, 24 bits of the rebased keys give the
        // same order as the 32-bit keys: three radix passes, not four.
        KeyRebase rb;
        const bool fit = depth_rebase(*view, sp->bmin, sp->bmax, &rb.base);
        uint2* ordered = nullptr;
        rc = argsort_payload_u32(c, s->sort_depth, s->depth_keys, s->cbbox, n, 1, fit ? 24 : 32,
                                 &ordered, fit ? &rb : nullptr);
        if (rc) return rc;
        if (binned) {
            // Each tile's isects straight from the depth order, no isect sort.
            BinJob j;
            memset(&j, 0, sizeof j);
            j.ordered = ordered;
            j.n = n;
            j.view0.tiles_x = s->tiles_x;
            j.view0.tiles_y = s->tiles_y;
            j.view0.supers_x = div_up(s->tiles_x, kBinSuper);
            j.view0.super_shift = kBinSuperShift;
            j.view0.out_mul = 1;
            j.nviews = 1;
            j.supers = bin_supers(s->tiles_x, s->tiles_y);
            j.view_shift = std::max(1u, sfk_bits_for(j.supers));
            j.out_tiles = tiles;
            j.cap = s->capacity;
            j.order_out = s->order_buf;
            j.offsets = s->offsets;
            j.offs_by_splat = s->bin_offs;
            j.sort = s->sort_tile;
            j.sstart = s->bin_sstart;
            j.chunk_map = s->bin_map;
            j.chunk_counts = s->bin_counts;
            j.tile_starts = s->tile_starts;
            j.tile_order = s->tile_order;
            j.gids = s->gaussian_ids;
            j.tile_ids = s->want_tile_ids && !s->pair_raster ? s->tile_ids : nullptr;
            j.skip_place = s->pair_raster;
            j.pairs_out = &s->pairs;
            j.class_count = s->class_count;
            j.tile_cls = s->tile_cls;
            j.counters = ctr;
            j.status = s->emit_status;
            j.status_zeroed = true;
            j.self_clean = true;
            j.clear = s->counters + (size_t)(bank ^ 1u) * kCounterWords + 2;
            j.clear_n = 2 * kCountSlots;
            j.host_total = s->host_counters_dev;  // the raw count, for next frame's capacity
            rc = bin_emit(c, j);
            if (rc) return rc;
            s->pairs_ready = s->pair_raster;
            s->bank = bank ^ 1u;
            s->emit_clean = true;
            s->order_valid = s->tile_order != nullptr;
            SFK_HIP(c, hipEventRecord(s->counters_ready, c->stream));
            s->counters_pending = true;
            s->depth_order = s->order_buf;
            s->sorted_tiles = j.tile_ids;
            s->sorted_gids = s->gaussian_ids;
            s->view = *view;
            s->has_view = true;
            fill_isects(s, out);
            return SFK_OK;
        }
        src.ordered = ordered;
        src.order_out = s->order_buf;
        s->depth_order = s->order_buf;
    } else {
        uint32_t* order = nullptr;
        rc = sfk_argsort_u32(c, s->sort_depth, s->depth_keys, nullptr, n, nullptr, 32, 0,
                             nullptr, &order);
        if (rc) return rc;
        src.order = order;
        src.tile_counts = s->tile_counts;
        src.bbox = s->bbox;
        s->depth_order = order;
    }
    {
        // Emission offsets in depth order and the isects themselves, in one
        // pass; the grand total is the raw isect count (counters[0]).
        SFK_PROF(c, "scan_emit");
        SFK_HIP(c, launch_scan_emit(c->stream, src, n, s->tiles_x, s->capacity, s->tile_ids,
                                    s->gaussian_ids, s->offsets, ctr, s->emit_status));
    }
    SFK_HIP(c, hipMemcpyAsync(s->host_counters, ctr, 4, hipMemcpyDeviceToHost, c->stream));
    SFK_HIP(c, hipEventRecord(s->counters_ready, c->stream));
    s->counters_pending = true;
    uint32_t* st = nullptr;
    uint32_t* sg = nullptr;
    rc = sfk_argsort_u32(c, s->sort_tile, s->tile_ids, s->gaussian_ids, s->capacity, ctr,
                         sfk_bits_for(tiles), 1, &st, &sg);
    if (rc) return rc;
    s->sorted_tiles = st;
    s->sorted_gids = sg;
    {
        SFK_PROF(c, "tile_starts");
        SFK_HIP(c, launch_tile_starts(c->stream, st, s->capacity, ctr, tiles,
                                      s->tile_starts));
    }
    s->view = *view;
    s->has_view = true;
    fill_isects(s, out);
    return SFK_OK;
}

extern "C" int sfk_prepare_isects(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                                  const sfk_camera* cam, uint32_t w, uint32_t h,
                                  sfk_tile_isects* out) {
    if (!c || !s || !cam) return SFK_E_ARG;
    sfk_view v;
    int rc = sfk_camera_view(cam, w, h, s->tile_w, &v);
    if (rc) return fail(c, rc, "camera view: bad arguments");
    return sfk_prepare_isects_view(c, sp, s, &v, w, h, out);
}

extern "C" int sfk_tile_isects_counts(sfk_ctx* c, const sfk_tile_isects* t, uint32_t counts[3]) {
    if (!c || !t || !counts) return SFK_E_ARG;
    uint32_t h[kCounterWords];
    SFK_HIP(c, hipMemcpyAsync(h, t->d_counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    uint32_t vis = 0;
    for (uint32_t k = 0; k < kCountSlots; ++k) vis += h[2 + 2 * k + 1];
    counts[0] = h[0];
    counts[1] = vis;
    counts[2] = std::min(h[0], t->capacity);
    return SFK_OK;
}

static RasterArgs raster_args(sfk_render_scratch* s, const sfk_tile_isects* t) {
    RasterArgs a;
    memset(&a, 0, sizeof a);
    a.w = t->w; a.h = t->h; a.tiles_x = t->tiles_x; a.tiles_y = t->tiles_y;
    a.tile_w = t->tile_w; a.row_stride = t->row_stride;
    a.tile_starts = t->tile_starts;
    a.tile_order = s->order_valid ? s->tile_order : nullptr;
    a.gids = t->gaussian_ids;
    a.proj = s->projected;
    if (s->pairs_ready) {
        a.pairs = s->pairs;
        a.pair_starts = s->bin_sstart;
        a.offs_by_splat = s->bin_offs;
        a.cap = s->capacity;
        a.gids_out = s->gaussian_ids;
    }
    return a;
}

static int check_isects(sfk_ctx* c, sfk_render_scratch* s, const sfk_tile_isects* t) {
    if (!c || !s || !t) return SFK_E_ARG;
    if (t->w != s->w || t->h != s->h || t->tile_w != s->tile_w || !t->gaussian_ids)
        return fail(c, SFK_E_ARG, "isects were not prepared by this scratch");
    return SFK_OK;
}

extern "C" int sfk_finalize_rgb(sfk_ctx* c, sfk_render_scratch* s, const sfk_tile_isects* t) {
    int rc = check_isects(c, s, t);
    if (rc) return rc;
    if (!s->has_colors)
        return fail(c, SFK_E_ARG, "finalize(Rgb) after a responsibility-only prepare");
    RasterArgs a = raster_args(s, t);
    a.rgba = s->rgba;
    a.bitmap = s->bitmap;
    a.final_t = s->final_t;
    a.n_contrib = s->n_contrib;
    SFK_PROF(c, "raster_rgb");
    SFK_HIP(c, launch_rasterize(c->stream, kRgb, a));
    return SFK_OK;
}

// The accumulators' fixed-point scale: 2^k for k = fractional_bits(w, h) in
// [0, 32] (seg/mod.rs:31-34, 178).  The rasterizers rely on it being a power
// of two (they carry T pre-scaled, exactly).
static bool fixed_point_scale(float scale) {
    const uint32_t u = __builtin_bit_cast(uint32_t, scale);
    return (u & 0x807fffffu) == 0 && scale >= 1.0f && scale <= 4294967296.0f;
}

extern "C" int sfk_finalize_total(sfk_ctx* c, sfk_render_scratch* s, const sfk_tile_isects* t,
                                  float scale, uint32_t* resp) {
    int rc = check_isects(c, s, t);
    if (rc) return rc;
    if (!resp) return fail(c, SFK_E_ARG, "finalize_total: null resp");
    if (!fixed_point_scale(scale))
        return fail(c, SFK_E_ARG, "finalize_total: scale %g is not 2^k, k in [0, 32]", scale);
    RasterArgs a = raster_args(s, t);
    a.scale = scale;
    a.resp = resp;
    SFK_PROF(c, "raster_total");
    SFK_HIP(c, launch_rasterize(c->stream, kTotal, a));
    return SFK_OK;
}

extern "C" int sfk_finalize_evidence(sfk_ctx* c, sfk_render_scratch* s, const sfk_tile_isects* t,
                                     float scale, const uint32_t* mask, uint32_t words_per_row,
                                     uint32_t* fg, uint32_t* bg) {
    int rc = check_isects(c, s, t);
    if (rc) return rc;
    if (!mask || !fg || !bg || words_per_row < (t->w + 31) / 32)
        return fail(c, SFK_E_ARG, "finalize_evidence: bad mask/fg/bg");
    if (!fixed_point_scale(scale))
        return fail(c, SFK_E_ARG, "finalize_evidence: scale %g is not 2^k, k in [0, 32]", scale);
    RasterArgs a = raster_args(s, t);
    a.scale = scale;
    a.mask = mask;
    a.words_per_row = words_per_row;
    a.fg = fg;
    a.bg = bg;
    SFK_PROF(c, "raster_evidence");
    SFK_HIP(c, launch_rasterize(c->stream, kEvidence, a));
    return SFK_OK;
}

extern "C" int sfk_render_with(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                               const sfk_camera* cam, uint32_t w, uint32_t h,
                               const uint32_t** bitmap, const float** rgba) {
    sfk_tile_isects t;
    int rc = sfk_prepare_isects(c, sp, s, cam, w, h, &t);
    if (rc) return rc;
    rc = sfk_finalize_rgb(c, s, &t);
    if (rc) return rc;
    if (bitmap) *bitmap = s->bitmap;
    if (rgba) *rgba = s->rgba;
    return SFK_OK;
}

// ---------------------------------------------------------------------- seg

extern "C" uint32_t sfk_fractional_bits(uint32_t w, uint32_t h) {
    uint64_t p = (uint64_t)w * h, np2 = 1;
    uint32_t tz = 0;
    while (np2 < p) { np2 <<= 1; ++tz; }
    return tz >= 32 ? 0 : 32 - tz;
}

extern "C" int sfk_mask_pack(const uint8_t* bytes, uint32_t w, uint32_t h, uint32_t* words,
                             uint32_t* fg_pixels) {
    if (!bytes || !words || w == 0) return SFK_E_ARG;
    uint32_t wpr = (w + 31) / 32;
    memset(words, 0, (size_t)wpr * h * 4);
    uint32_t fg = 0;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            if (bytes[(size_t)y * w + x]) {
                words[(size_t)y * wpr + x / 32] |= 1u << (x % 32);
                ++fg;
            }
    if (fg_pixels) *fg_pixels = fg;
    return SFK_OK;
}

static float scale_for(uint32_t w, uint32_t h) {
    return (float)(1ull << sfk_fractional_bits(w, h));
}

extern "C" int sfk_accumulate_prepared(sfk_ctx* c, sfk_render_scratch* s,
                                       const sfk_tile_isects* t, const uint32_t* mask,
                                       uint32_t words_per_row, uint32_t* resp, uint32_t* fg,
                                       uint32_t* bg, float* scale_out) {
    int rc = check_isects(c, s, t);
    if (rc) return rc;
    float scale = scale_for(t->w, t->h);
    if (scale_out) *scale_out = scale;
    const uint32_t n = s->n;
    if (!mask) {
        if (!resp) return fail(c, SFK_E_ARG, "accumulate: null resp");
        SFK_HIP(c, launch_fill_u32(c->stream, resp, 0u, n));
        return sfk_finalize_total(c, s, t, scale, resp);
    }
    if (!fg || !bg) return fail(c, SFK_E_ARG, "accumulate: null fg/bg");
    SFK_HIP(c, launch_fill_u32(c->stream, fg, 0u, n));
    SFK_HIP(c, launch_fill_u32(c->stream, bg, 0u, n));
    return sfk_finalize_evidence(c, s, t, scale, mask, words_per_row, fg, bg);
}

extern "C" int sfk_render_responsibility(sfk_ctx* c, const sfk_splats* sp,
                                         sfk_render_scratch* s, const sfk_camera* cam, uint32_t w,
                                         uint32_t h, uint32_t* resp, float* scale_out) {
    if (!c || !s || !cam) return SFK_E_ARG;
    sfk_view v;
    sfk_tile_isects t;
    int rc = sfk_camera_view(cam, w, h, s->tile_w, &v);
    if (rc) return fail(c, rc, "camera view: bad arguments");
    rc = prepare_impl(c, sp, s, &v, w, h, &t, false);
    if (rc) return rc;
    return sfk_accumulate_prepared(c, s, &t, nullptr, 0, resp, nullptr, nullptr, scale_out);
}

// At least `np` doubles of context workspace.
static int workspace_doubles(sfk_ctx* c, uint32_t np) {
    if (np > c->eig_partials_cap) {
        if (c->eig_partials) hipFree(c->eig_partials);
        c->eig_partials = nullptr;
        c->eig_partials_cap = 0;
        SFK_HIP(c, hipMalloc(&c->eig_partials, (size_t)np * 8));
        c->eig_partials_cap = np;
    }
    return SFK_OK;
}

static int eig_workspace(sfk_ctx* c, uint32_t n) { return workspace_doubles(c, eig_num_partials(n)); }

extern "C" int sfk_eig_device(sfk_ctx* c, const float* a, const float* b, const uint32_t* resp,
                              uint32_t n, float scale, double* d_out) {
    if (!c || !d_out || (n && (!a || !b || !resp))) return SFK_E_ARG;
    int rc = eig_workspace(c, n);
    if (rc) return rc;
    SFK_PROF(c, "eig");
    SFK_HIP(c, launch_eig(c->stream, a, b, resp, n, scale, c->eig_partials, d_out));
    return SFK_OK;
}

extern "C" int sfk_eig(sfk_ctx* c, const float* a, const float* b, const uint32_t* resp,
                       uint32_t n, float scale, double* out) {
    if (!c || !out) return SFK_E_ARG;
    int rc = sfk_eig_device(c, a, b, resp, n, scale, c->eig_out);
    if (rc) return rc;
    SFK_HIP(c, hipMemcpyAsync(out, c->eig_out, 8, hipMemcpyDeviceToHost, c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    return SFK_OK;
}

extern "C" int sfk_beta_update(sfk_ctx* c, float* a, float* b, const uint32_t* fg,
                               const uint32_t* bg, uint32_t n, float scale) {
    if (!c || (n && (!a || !b || !fg || !bg))) return SFK_E_ARG;
    SFK_HIP(c, launch_beta_update(c->stream, a, b, fg, bg, n, scale));
    return SFK_OK;
}

// -------------------------------------------------------------- paint

extern "C" int sfk_render_posterior(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                                    const float* a, const float* b, uint32_t n_ab,
                                    const sfk_camera* cam, uint32_t w, uint32_t h,
                                    const uint32_t** bitmap, const float** rgba) {
    if (!c || !sp || !s) return SFK_E_ARG;
    // A posterior sized for another splat count renders the normal colours
    // (paint.rs:30-32).
    if (!a || !b || n_ab != sp->n) return sfk_render_with(c, sp, s, cam, w, h, bitmap, rgba);
    if (!cam) return SFK_E_ARG;
    sfk_view v;
    int rc = sfk_camera_view(cam, w, h, s->tile_w, &v);
    if (rc) return fail(c, rc, "camera view: bad arguments");
    // The paint rides in the projection: rows get m = a/(a+b) and no SH is
    // read (paint.rs projects SH colours, then overwrites them).
    sfk_tile_isects t;
    rc = prepare_impl(c, sp, s, &v, w, h, &t, false, a, b);
    if (rc) return rc;
    rc = sfk_finalize_rgb(c, s, &t);
    if (rc) return rc;
    if (bitmap) *bitmap = s->bitmap;
    if (rgba) *rgba = s->rgba;
    return SFK_OK;
}

extern "C" int sfk_tint(sfk_ctx* c, sfk_splats* sp, const float* a, const float* b,
                        uint32_t n_ab, const float color[3]) {
    if (!c || !sp || !a || !b || !color) return SFK_E_ARG;
    if (n_ab != sp->n) return fail(c, SFK_E_ARG, "tint: %u posteriors for %u splats", n_ab, sp->n);
    SFK_HIP(c, launch_tint_colors(c->stream, sp->sh, a, b, sp->n, color[0], color[1], color[2]));
    return SFK_OK;
}

extern "C" int sfk_save_colors(sfk_ctx* c, const sfk_splats* sp, float* saved) {
    if (!c || !sp || !saved) return SFK_E_ARG;
    SFK_HIP(c, launch_copy_f32(c->stream, saved, sp->sh, (size_t)3 * sp->n));
    return SFK_OK;
}

extern "C" int sfk_restore_colors(sfk_ctx* c, sfk_splats* sp, const float* saved) {
    if (!c || !sp || !saved) return SFK_E_ARG;
    SFK_HIP(c, launch_copy_f32(c->stream, sp->sh, saved, (size_t)3 * sp->n));
    return SFK_OK;
}

// ----------------------------------------------------------- kernel-level

extern "C" int sfk_project(sfk_ctx* c, const sfk_view* v, const float* attrs, const float* sh,
                           uint32_t n, uint32_t sh_per_ch, uint32_t tile_w, uint32_t* depth_keys,
                           float* projected, uint32_t* tile_counts, uint32_t* bbox,
                           uint32_t* counters) {
    if (!c || !v || (n && (!attrs || !depth_keys || !projected || !tile_counts || !bbox || !counters)))
        return SFK_E_ARG;
    if (sh_per_ch && !sh) return SFK_E_ARG;
    if (!(tile_w == 8 || tile_w == 16 || tile_w == 32)) return fail(c, SFK_E_ARG, "tile width %u", tile_w);
    SFK_HIP(c, launch_project(c->stream, *v, attrs, sh, n, sh_per_ch, tile_w, depth_keys, projected,
                              tile_counts, bbox, counters));
    return SFK_OK;
}

extern "C" int sfk_map_isects(sfk_ctx* c, const uint32_t* order, const uint32_t* tile_counts,
                              const uint32_t* bbox, const uint32_t* offsets, uint32_t n,
                              uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                              uint32_t* gaussian_ids) {
    if (!c || (n && (!order || !tile_counts || !bbox || !offsets || !tile_ids || !gaussian_ids)))
        return SFK_E_ARG;
    SFK_HIP(c, launch_map_isects(c->stream, order, tile_counts, bbox, offsets, n, tiles_per_row,
                                 max_isects, tile_ids, gaussian_ids));
    return SFK_OK;
}

extern "C" int sfk_tile_starts(sfk_ctx* c, const uint32_t* tile_ids, uint32_t n,
                               const uint32_t* d_n, uint32_t num_tiles, uint32_t* starts) {
    if (!c || !starts || (n && !tile_ids)) return SFK_E_ARG;
    SFK_HIP(c, launch_tile_starts(c->stream, tile_ids, n, d_n, num_tiles, starts));
    return SFK_OK;
}

// ----------------------------------------------------------------- backward

extern "C" int sfk_loss_l2(sfk_ctx* c, const float* rgba, const float* target, uint32_t pixels,
                           float* d_rgba, double* d_loss) {
    if (!c || (pixels && (!rgba || !target || !d_rgba))) return SFK_E_ARG;
    int rc = workspace_doubles(c, loss_num_partials(pixels));
    if (rc) return rc;
    SFK_HIP(c, launch_loss_l2(c->stream, rgba, target, pixels, d_rgba, c->eig_partials, d_loss));
    return SFK_OK;
}

static int render_backward_impl(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                                const sfk_tile_isects* t, const float* d_rgba,
                                const float* rgba, const float* target, float* grad_attrs,
                                float* grad_sh, double* d_loss) {
    int rc = check_isects(c, s, t);
    if (rc) return rc;
    if (!sp || !grad_attrs || !grad_sh || (!d_rgba && (!rgba || !target || !d_loss)))
        return fail(c, SFK_E_ARG, "backward: null buffer");
    if (sp->n != s->n || !s->has_view) return fail(c, SFK_E_ARG, "backward: scratch was not prepared for these splats");
    // The adjoint differentiates the SH colour; rows that hold posterior
    // means (render_posterior) or no colour (a responsibility-only prepare)
    // would give silently wrong gradients.
    if (!s->sh_colors)
        return fail(c, SFK_E_ARG, "backward: the last prepare did not project SH colours");
    if (!s->grad2d) SFK_HIP(c, hipMalloc(&s->grad2d, (size_t)s->n * 9 * 4));
    // The projection adjoint clears what it reads, so only a fresh buffer (or
    // one left by an interrupted call) needs the fill.
    if (s->grad2d_dirty)
        SFK_HIP(c, launch_fill_u32(c->stream, reinterpret_cast<uint32_t*>(s->grad2d), 0u,
                                   (size_t)s->n * 9));
    s->grad2d_dirty = true;
    RasterArgs a = raster_args(s, t);
    BwdLoss loss;
    if (!d_rgba) {
        const uint32_t tiles = s->tiles_x * s->tiles_y;
        if (s->loss_partials_cap < tiles) {
            if (s->loss_partials) hipFree(s->loss_partials);
            s->loss_partials = nullptr;
            s->loss_partials_cap = 0;
            SFK_HIP(c, hipMalloc(&s->loss_partials, (size_t)tiles * 8));
            s->loss_partials_cap = tiles;
        }
        loss = {rgba, target, s->loss_partials, tiles, d_loss};
    }
    {
        SFK_PROF(c, "raster_bwd");
        SFK_HIP(c, launch_raster_bwd(c->stream, a, s->final_t, s->n_contrib, d_rgba, s->n,
                                     s->grad2d, d_rgba ? nullptr : &loss));
    }
    SFK_PROF(c, "project_bwd");
    SFK_HIP(c, launch_project_bwd(c->stream, s->view, sp->attrs, sp->sh, sp->n, sp->sh_per_ch,
                                  s->grad2d, grad_attrs, grad_sh, d_rgba ? nullptr : &loss,
                                  s->has_dcol ? s->dcol : nullptr));
    s->grad2d_dirty = false;
    // From the next frame on, the forward keeps d colour / d direction.
    if (!s->dcol && sp->sh_per_ch > 1) SFK_HIP(c, hipMalloc(&s->dcol, (size_t)s->n * 9 * 4));
    s->want_dcol = s->dcol != nullptr;
    return SFK_OK;
}

extern "C" int sfk_localize(sfk_ctx* c, const sfk_splats* sp, const float* a, const float* b,
                            float* center3, float* radius, int* found) {
    if (!c || !sp || !a || !b || !center3 || !radius || !found) return SFK_E_ARG;
    const uint32_t nb = localize_blocks(sp->n);
    int rc = workspace_doubles(c, nb * 4u + 8u);
    if (rc) return rc;
    double* out = c->eig_partials + (size_t)nb * 4;
    SFK_HIP(c, launch_localize(c->stream, sp->attrs, a, b, sp->n, c->eig_partials, out));
    double h[5];
    SFK_HIP(c, hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    *found = h[3] > 0.0 ? 1 : 0;
    center3[0] = (float)h[0];
    center3[1] = (float)h[1];
    center3[2] = (float)h[2];
    *radius = (float)h[4];
    return SFK_OK;
}

extern "C" int sfk_render_backward(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                                   const sfk_tile_isects* t, const float* d_rgba,
                                   float* grad_attrs, float* grad_sh) {
    if (!d_rgba) return c ? fail(c, SFK_E_ARG, "backward: null d_rgba") : SFK_E_ARG;
    return render_backward_impl(c, sp, s, t, d_rgba, nullptr, nullptr, grad_attrs, grad_sh,
                                nullptr);
}

extern "C" int sfk_render_backward_l2(sfk_ctx* c, const sfk_splats* sp, sfk_render_scratch* s,
                                      const sfk_tile_isects* t, const float* rgba,
                                      const float* target, float* grad_attrs, float* grad_sh,
                                      double* d_loss) {
    return render_backward_impl(c, sp, s, t, nullptr, rgba, target, grad_attrs, grad_sh, d_loss);
}
