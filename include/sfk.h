/* sfk.h — C ABI of libsfk, the MI355X (gfx950) implementation of the
 * splatfield 3DGS render path and B3-Seg evidence path.
 *
 * Every entry point enqueues on the context's single HIP stream (the ordering
 * contract of src/lib.rs:20-27) unless marked SYNC.  Return codes: 0 = ok,
 * negative = error; sfk_last_error() gives the message.  No exception or abort
 * crosses this boundary: the reference's panics (undersized scratch,
 * sort.rs:323-327, scan.rs:230-234; zero splats, render.rs:248) become
 * SFK_E_CAPACITY / SFK_E_ARG.  Calls on one context must be externally
 * serialized (the reference takes &mut for the same reason).
 *
 * Device pointers are plain pointers from sfk_alloc (or any hipMalloc'd
 * memory of the same device).  Objects returned through out-parameters alias
 * their scratch and stay valid until the next call on that scratch
 * (src/render.rs:366-370; crates/sort/src/sort.rs:278-281).
 */
#ifndef SFK_H
#define SFK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFK_OK 0
#define SFK_E_ARG (-1)
#define SFK_E_CAPACITY (-2)
#define SFK_E_HIP (-3)
#define SFK_E_NOMEM (-4)

#define SFK_ATTR_PLANES 11   /* src/layout.rs:47-58 */
#define SFK_PROJ_STRIDE 12   /* floats per projected row (see sfk_render_buffers) */

/* ------------------------------------------------------------ context / L0 */

typedef struct sfk_ctx sfk_ctx;

/* One device + one HIP stream; replaces the cubecl client and its
 * single-stream policy (src/lib.rs:20-27, crates/sort/src/tensor.rs:15-27). */
int sfk_ctx_create(int device, sfk_ctx** out);
void sfk_ctx_destroy(sfk_ctx* ctx);
const char* sfk_last_error(const sfk_ctx* ctx);
void* sfk_ctx_stream(sfk_ctx* ctx); /* the hipStream_t */
int sfk_sync(sfk_ctx* ctx);         /* SYNC */
int sfk_device_name(sfk_ctx* ctx, char* buf, int len);

/* GpuTensor storage (crates/sort/src/tensor.rs:29-77): empty / from /
 * read_vec / write.  Upload and download are SYNC. */
int sfk_alloc(sfk_ctx* ctx, size_t bytes, void** out);
int sfk_free(sfk_ctx* ctx, void* ptr);
int sfk_upload(sfk_ctx* ctx, void* dst, const void* src, size_t bytes);
int sfk_download(sfk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Stream-ordered u32 fill by a kernel (graph-capture safe); zero_u32 of
 * src/seg/mod.rs:87-103 is fill(buf, 0). */
int sfk_fill_u32(sfk_ctx* ctx, uint32_t* buf, uint32_t value, size_t count);
/* Stream-ordered device copy of count u32 (count a multiple of 4). */
int sfk_copy_u32(sfk_ctx* ctx, uint32_t* dst, const uint32_t* src, size_t count);

/* Stream-ordered event timing (hipEventRecord/ElapsedTime). */
int sfk_timer_start(sfk_ctx* ctx);
int sfk_timer_stop_ms(sfk_ctx* ctx, float* ms); /* SYNC */
/* Per-launch timing: with profiling on, every kernel-launching step records
 * an event pair on the stream; the report (SYNC) aggregates them by name as
 * "name count total_ms" lines and clears the record. */
int sfk_prof_enable(sfk_ctx* ctx, int on);
int sfk_prof_report(sfk_ctx* ctx, char* buf, int len);

/* ------------------------------------------------------------------ camera */

/* Camera (src/camera.rs:4-19): per-axis fov, position, rotation (x,y,z,w). */
typedef struct {
    float fov[2];
    float position[3];
    float rotation[4];
} sfk_camera;

/* Kernel view of a camera (CameraView, src/project.rs:60-74). */
typedef struct {
    float rot0[3], rot1[3], rot2[3]; /* rows of the world->camera rotation */
    float trans[3];
    float focal[2];
    float world_pos[3];
    float img[2];
    float tile_bounds[2];
} sfk_view;

/* CameraViewLaunch::for_camera (src/project.rs:76-96). */
int sfk_camera_view(const sfk_camera* cam, uint32_t w, uint32_t h, uint32_t tile_w,
                    sfk_view* out);
/* Camera::look_at rotation (src/camera.rs:57-63). */
int sfk_look_at(const float pos[3], const float target[3], float quat_xyzw[4]);

/* -------------------------------------------------------- crates/sort (L1) */

/* bits_for (crates/sort/src/sort.rs:259-261). */
uint32_t sfk_bits_for(uint32_t max_exclusive);

typedef struct sfk_radix_scratch sfk_radix_scratch;

/* RadixScratch::new (sort.rs:290-300).  key_bytes is 4 (u32 keys) or 8
 * (u64 keys). */
int sfk_radix_scratch_create(sfk_ctx* ctx, uint32_t max_elems, uint32_t key_bytes,
                             sfk_radix_scratch** out);
void sfk_radix_scratch_destroy(sfk_radix_scratch* s);
/* Digit width per pass, 4..8 bits (default 8).  The reference fixes 6
 * (sort.rs:11-17); here it is a tunable. */
int sfk_radix_set_digit_bits(sfk_radix_scratch* s, uint32_t digit_bits);
/* Strategy of u64 sorts (a tuning and test knob, no counterpart in the
 * reference; on by default): a sort of at least 2^20 keys of >= 24 bits
 * with a host count runs two stable passes on the keys' top 16 bits, then
 * sorts every resulting group of at most 384 keys by the remaining bits on
 * chip in one launch, and reads back the largest group size (the call
 * synchronises the context stream once); if a group was larger, it runs the
 * plain passes instead.  Same results either way. */
int sfk_radix_set_bucket_finish(sfk_radix_scratch* s, int on);

/* RadixScratch::argsort (sort.rs:302-381): stable argsort of keys[0..n) on
 * their low `bits` bits (keys must be < 2^bits), carrying vals.  vals may be
 * NULL: the identity 0..n-1 is used.  If d_n is non-NULL the element count is
 * min(*d_n, n) read on the device (no host sync).  With write_keys = 0 the
 * sorted keys of the last pass are not written.  Inputs are only read;
 * outputs alias the scratch until its next use. */
int sfk_argsort_u32(sfk_ctx* ctx, sfk_radix_scratch* s, const uint32_t* keys,
                    const uint32_t* vals, uint32_t n, const uint32_t* d_n, uint32_t bits,
                    int write_keys, uint32_t** out_keys, uint32_t** out_vals);
int sfk_argsort_u64(sfk_ctx* ctx, sfk_radix_scratch* s, const uint64_t* keys,
                    const uint32_t* vals, uint32_t n, const uint32_t* d_n, uint32_t bits,
                    int write_keys, uint64_t** out_keys, uint32_t** out_vals);

/* nseg independent stable argsorts of seg_len keys each, laid out back to
 * back, in the same launches (no reference counterpart: the batched
 * multi-view depth sort).  Identity values are indices within the segment. */
int sfk_argsort_u32_segmented(sfk_ctx* ctx, sfk_radix_scratch* s, const uint32_t* keys,
                              const uint32_t* vals, uint32_t nseg, uint32_t seg_len,
                              uint32_t bits, int write_keys, uint32_t** out_keys,
                              uint32_t** out_vals);

typedef struct sfk_scan_scratch sfk_scan_scratch;

/* ScanScratch::new (crates/sort/src/scan.rs:53-60). */
int sfk_scan_scratch_create(sfk_ctx* ctx, uint32_t max_elems, sfk_scan_scratch** out);
void sfk_scan_scratch_destroy(sfk_scan_scratch* s);
/* exclusive_scan_gather (scan.rs:215-255): offsets[i] = sum_{j<i}
 * counts[gids[j]] for i < n; n = 0 leaves offsets untouched. */
int sfk_exclusive_scan_gather(sfk_ctx* ctx, const uint32_t* gids, const uint32_t* counts,
                              uint32_t* offsets, uint32_t n, sfk_scan_scratch* s);
/* exclusive_scan_buf (scan.rs:182-213): in-place exclusive scan of buf[0..n). */
int sfk_exclusive_scan(sfk_ctx* ctx, uint32_t* buf, uint32_t n, sfk_scan_scratch* s);

/* ---------------------------------------------------------- render (L2) */

typedef struct sfk_splats sfk_splats;

/* Splats::new (src/render.rs:241-269): upload field-major attribute planes
 * (11*n floats) and SH planes (3*sh_per_ch*n floats).  n = 0 is SFK_E_ARG. */
int sfk_splats_create(sfk_ctx* ctx, const float* attrs, const float* sh, uint32_t n,
                      uint32_t sh_per_ch, sfk_splats** out);
void sfk_splats_destroy(sfk_splats* s);
int sfk_splats_info(const sfk_splats* s, uint32_t* n, uint32_t* sh_per_ch, float bounds_min[3],
                    float bounds_max[3], const float** d_attrs, const float** d_sh);

typedef struct sfk_render_scratch sfk_render_scratch;

/* RenderScratch::new (src/render.rs:143-167): per-(n, w x h) buffers reused
 * across frames.  Isect capacity starts at next_pow2(min(tiles*n, 2^22)) and
 * grows between frames from the previous frame's emission up to 67,107,840
 * (render.rs:12-18, 169-197). */
int sfk_render_scratch_create(sfk_ctx* ctx, uint32_t n, uint32_t w, uint32_t h,
                              sfk_render_scratch** out);
void sfk_render_scratch_destroy(sfk_render_scratch* s);
/* Tile width 8, 16 (the reference's TILE_WIDTH, layout.rs:41) or 32. */
int sfk_render_scratch_set_tile_width(sfk_render_scratch* s, uint32_t tile_w);
/* Raise the isect capacity now (next_pow2, capped); a pre-sized scratch
 * renders its first frame untruncated. */
int sfk_render_scratch_reserve_isects(sfk_render_scratch* s, uint32_t isects);
/* Radix digit width of the scratch's sorts (4..8). */
int sfk_render_scratch_set_digit_bits(sfk_render_scratch* s, uint32_t digit_bits);
/* Isect stage on grids of at most 255 x 255 tiles: on (the default), the
 * binned emission writes every tile's isects in depth order without sorting
 * them; off, emission + stable tile argsort as render.rs:271-364 does.  Both
 * give the same arrays word for word (a tuning and test knob, no counterpart
 * in the reference). */
int sfk_render_scratch_set_binned(sfk_render_scratch* s, int on);
/* Whether the binned emission also writes the per-isect tile ids (on, the
 * default).  The rasterizers read only tile_starts and gaussian_ids, so a
 * caller that never reads tile_ids (a training loop) switches them off and
 * saves one 4-byte store per isect; sfk_tile_isects.tile_ids is then NULL.
 * The sorting emission always has them (they are its sort keys).  No
 * counterpart in the reference, whose emission sorts by tile id. */
int sfk_render_scratch_set_tile_ids(sfk_render_scratch* s, int on);
/* Pair raster (off by default; binned emission only).  The emission stops
 * after the tile counts, tile_starts and the tile order, without placing
 * the isects; each rasterizer (finalize in any mode) takes a tile's isects
 * straight from its supertile's depth-ordered (splat, bbox) pairs, and
 * writes the ones it consumes -- every isect up to the tile's last
 * contributor -- into gaussian_ids at their slots.  Images, final T,
 * n_contrib, evidence words and the backward pass are unchanged;
 * sfk_tile_isects.gaussian_ids then holds only those consumed isects and
 * tile_ids is NULL.  For a caller that never reads the isect arrays (a
 * training loop).  No counterpart in the reference. */
int sfk_render_scratch_set_pair_raster(sfk_render_scratch* s, int on);

/* TileIsects (src/render.rs:463-471) plus the device-side counters. */
typedef struct {
    const uint32_t* tile_ids;     /* sorted, tile-major / depth-minor */
    const uint32_t* gaussian_ids; /* splat id per isect */
    const uint32_t* tile_starts;  /* lower_bound(tile) for tile in [0, tiles] */
    const uint32_t* d_counters;   /* device [isects_raw, visible, isects] */
    uint32_t capacity;
    uint32_t tiles_x, tiles_y, tile_w, row_stride;
    uint32_t w, h;
} sfk_tile_isects;

/* Splats::prepare_isects (src/render.rs:271-364): project, depth argsort,
 * gather scan, map, stable tile argsort, tile ranges.  No host sync: counts
 * stay on the device (sfk_tile_isects_counts reads them). */
int sfk_prepare_isects(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                       const sfk_camera* cam, uint32_t w, uint32_t h, sfk_tile_isects* out);
/* The same with an explicit kernel view (tests and batched callers). */
int sfk_prepare_isects_view(sfk_ctx* ctx, const sfk_splats* splats,
                            sfk_render_scratch* scratch, const sfk_view* view, uint32_t w,
                            uint32_t h, sfk_tile_isects* out);
/* SYNC: [isects_raw, visible, isects]. */
int sfk_tile_isects_counts(sfk_ctx* ctx, const sfk_tile_isects* isects, uint32_t counts[3]);

/* Finalize (src/render.rs:220-239, 382-460) in its three modes.  The
 * responsibility modes take the accumulators' fixed-point scale, 2^k with
 * k = fractional_bits(w, h) in [0, 32] (src/seg/mod.rs:31-34); any other
 * scale is rejected (SFK_E_ARG). */
int sfk_finalize_rgb(sfk_ctx* ctx, sfk_render_scratch* scratch, const sfk_tile_isects* isects);
int sfk_finalize_total(sfk_ctx* ctx, sfk_render_scratch* scratch, const sfk_tile_isects* isects,
                       float scale, uint32_t* resp);
int sfk_finalize_evidence(sfk_ctx* ctx, sfk_render_scratch* scratch,
                          const sfk_tile_isects* isects, float scale, const uint32_t* mask,
                          uint32_t words_per_row, uint32_t* fg, uint32_t* bg);

/* Splats::render_with (src/render.rs:366-380): prepare + finalize(Rgb).
 * Outputs alias the scratch: bitmap (RGBA8, h x row_stride u32) and the
 * f32 RGBA image (h x w x 4). */
int sfk_render_with(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                    const sfk_camera* cam, uint32_t w, uint32_t h, const uint32_t** bitmap,
                    const float** rgba);

/* The scratch's device buffers, for inspection.  projected rows are
 * SFK_PROJ_STRIDE floats: [mx, my, conic_a, conic_b, conic_c, opacity, r, g,
 * b, 0, 0, 0] (the reference's 9-float row, src/layout.rs:60-61, reordered
 * and padded to 16-byte loads).  depth_keys[i] is the f32 bits of camera z
 * for visible splats and 0xffffffff otherwise (index i, not slot order);
 * depth_order[r] is the splat of depth rank r. */
typedef struct {
    const uint32_t* depth_keys;
    const uint32_t* depth_order;
    const float* projected;
    const uint32_t* tile_counts;
    const uint32_t* bbox;
    const uint32_t* offsets;
    const uint32_t* counters;
    const uint32_t* bitmap;
    const float* rgba;
    const float* final_t;
    const uint32_t* n_contrib;
    uint32_t n, w, h, row_stride, capacity;
} sfk_render_buffers;
int sfk_render_scratch_buffers(const sfk_render_scratch* s, sfk_render_buffers* out);

/* --------------------------------------------------------------- seg (L3) */

/* fractional_bits (src/seg/mod.rs:25-34). */
uint32_t sfk_fractional_bits(uint32_t w, uint32_t h);

/* Mask::from_bytes (src/seg/mod.rs:115-143), host side: one byte per pixel,
 * row-major, nonzero = inside -> ceil(w/32)*h LSB-first words. */
int sfk_mask_pack(const uint8_t* bytes, uint32_t w, uint32_t h, uint32_t* words,
                  uint32_t* fg_pixels);

/* Splats::render_responsibility (src/seg/mod.rs:146-159): prepare, clear
 * resp, accumulate total epsilon at scale 2^fractional_bits(w, h). */
int sfk_render_responsibility(sfk_ctx* ctx, const sfk_splats* splats,
                              sfk_render_scratch* scratch, const sfk_camera* cam, uint32_t w,
                              uint32_t h, uint32_t* resp, float* scale_out);
/* accumulate_prepared (src/seg/mod.rs:161-210): over isects already prepared
 * for this camera and resolution; mask NULL -> total into resp, else split
 * into fg/bg.  Buffers are cleared first. */
int sfk_accumulate_prepared(sfk_ctx* ctx, sfk_render_scratch* scratch,
                            const sfk_tile_isects* isects, const uint32_t* mask,
                            uint32_t words_per_row, uint32_t* resp, uint32_t* fg, uint32_t* bg,
                            float* scale_out);

/* BetaState::eig (src/seg/beta.rs:240-268) with the entropy in f64 on the
 * device instead of the 445x445 table.  SYNC (one f64 readback). */
int sfk_eig(sfk_ctx* ctx, const float* a, const float* b, const uint32_t* resp, uint32_t n,
            float scale, double* out);
/* The same, leaving the f64 result in device memory (no sync). */
int sfk_eig_device(sfk_ctx* ctx, const float* a, const float* b, const uint32_t* resp,
                   uint32_t n, float scale, double* d_out);
/* sfk_eig_device without any allocation: no counterpart in the reference,
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops).  The per-block partials go to a fixed file-scope
 * device array of the library (one per device, 512 KiB, held until the
 * process ends) instead of the context's grown workspace; n above
 * SFK_EIG_NOALLOC_MAX_SPLATS is SFK_E_CAPACITY.  All calls on one device
 * share that array: calls that overlap on one device (two contexts or
 * streams) race. */
#define SFK_EIG_NOALLOC_MAX_SPLATS (1u << 26)
int sfk_eig_device_noalloc(sfk_ctx* ctx, const float* a, const float* b, const uint32_t* resp,
                           uint32_t n, float scale, double* d_out);
/* BetaState::update (src/seg/beta.rs:270-286): a += fg/scale, b += bg/scale. */
int sfk_beta_update(sfk_ctx* ctx, float* a, float* b, const uint32_t* fg, const uint32_t* bg,
                    uint32_t n, float scale);
/* views::localize (src/seg/views.rs:32-67) on the device: centre = sum(m p) /
 * sum(m) and radius = sum(m |p - centre|) / sum(m) over splats with a > b,
 * m = a / (a + b), p = the splat mean, accumulated in f64 (SYNC: 40 bytes
 * read back instead of the reference's 8 B/splat a,b readback).
 * *found = 0 when no splat is foreground (centre and radius then 0). */
int sfk_localize(sfk_ctx* ctx, const sfk_splats* splats, const float* a, const float* b,
                 float* center3, float* radius, int* found);

/* ------------------------------------------ scene ingestion (src/ply.rs) */

#define SFK_PLY_MAX_REST 45    /* f_rest_* properties: SH degree 3 */
#define SFK_PLY_MAX_STRIDE 128 /* float properties per vertex */

/* Header of a binary little-endian 3DGS PLY (parse_ply, src/ply.rs:9-75). */
typedef struct {
    uint64_t n;           /* vertices */
    uint64_t data_offset; /* bytes before the first vertex row */
    uint32_t stride;      /* float properties per vertex row */
    uint32_t n_rest;      /* f_rest_* properties */
    uint32_t sh_planes;   /* n_rest + 3 SH planes written (ply.rs:75) */
    uint32_t sh_per_ch;   /* n_rest / 3 + 1 */
    /* columns of x y z scale_0..2 opacity rot_0..3 f_dc_0..2 */
    int32_t base[14];
    int32_t rest[SFK_PLY_MAX_REST]; /* f_rest_k columns, ascending k */
} sfk_ply_layout;

/* parse_ply's header pass and checks (ply.rs:13-75): format, vertex count,
 * float-only properties, the 14 named properties (the first missing one
 * reported), f_rest_k sorted by k, and that every row is in the buffer.
 * Host only; errors are SFK_E_ARG with the reference's message in err. */
int sfk_ply_parse_header(const void* bytes, size_t len, sfk_ply_layout* out, char* err,
                         int err_len);
/* parse_ply's row loop (ply.rs:78-106) on the device: rows (n x stride
 * floats, as in the file) to the 11 attribute planes (rotation normalised)
 * and sh_planes SH planes (DC, then f_rest_k channel-major in the file to
 * coefficient-interleaved planes).  rows must be 4-byte aligned (a device
 * copy of the file plus data_offset may not be: SFK_E_ARG); 16-byte aligned
 * rows take vector loads, others scalar ones. */
int sfk_ply_decode(sfk_ctx* ctx, const sfk_ply_layout* layout, const float* rows, float* attrs,
                   float* sh);
/* Splats from a PLY file image in host memory (lib.rs:45-50 + Splats::new):
 * header on the host, rows decoded on the device.  SH degree 0-3. */
int sfk_splats_from_ply(sfk_ctx* ctx, const void* bytes, size_t len, sfk_splats** out);

/* SOG archives (src/sog.rs:97-198).  The caller unzips meta.json and decodes
 * the WebP sheets to RGBA8 (one 4-byte texel per splat, row-major, at least
 * n texels each; the centroid sheet any size); the device decodes the
 * splats. */
typedef struct {
    uint64_t n;             /* meta.count */
    uint64_t centroid_px;   /* texels in the shN centroid sheet (0 without shN) */
    float mins[3], maxs[3]; /* meta.means */
    uint32_t sh_per_ch;     /* 1, or 4 / 9 / 16 for shN.bands 1 / 2 / 3 */
    uint32_t centroid_w;    /* centroid sheet width in texels */
    uint32_t palette_count; /* centroid_px / (sh_per_ch - 1) */
} sfk_sog_meta;
/* The decode loop of parse_sog (sog.rs:108-195) into field-major planes
 * (11 x n attrs, 3 sh_per_ch x n sh).  codebooks: 768 device floats (scales,
 * sh0, shN, 256 each).  status (2 device words): the first splat with a
 * rotation tag below 252, and the first whose shN label is >= palette_count
 * or whose palette leaves the sheet; 0xffffffff when none (the reference
 * fails the whole parse on either). */
int sfk_sog_decode(sfk_ctx* ctx, const sfk_sog_meta* meta, const uint8_t* means_lo,
                   const uint8_t* means_hi, const uint8_t* scales, const uint8_t* quats,
                   const uint8_t* sh0, const uint8_t* labels, const uint8_t* centroids,
                   const float* codebooks, float* attrs, float* sh, uint32_t* status);
/* Splats::new over planes already in device memory (copied; the caller's
 * buffers stay the caller's). */
int sfk_splats_create_device(sfk_ctx* ctx, const float* d_attrs, const float* d_sh, uint32_t n,
                             uint32_t sh_per_ch, sfk_splats** out);

/* ------------------------------------------------ paint (src/seg/paint.rs) */

/* Splats::render_posterior (src/seg/paint.rs:22-47): render_with with every
 * splat coloured by its posterior mean m = a/(a+b) as grayscale, no +0.5
 * offset (paint_posterior, paint.rs:99-109).  n_ab != the splat count
 * renders the normal colours instead, as the reference does. */
int sfk_render_posterior(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                         const float* a, const float* b, uint32_t n_ab, const sfk_camera* cam,
                         uint32_t w, uint32_t h, const uint32_t** bitmap, const float** rgba);
/* Splats::tint (paint.rs:49-65, tint_colors 113-122): DC planes of the
 * splats with a > b set to color, in place.  n_ab must equal the splat count
 * (SFK_E_ARG otherwise). */
int sfk_tint(sfk_ctx* ctx, sfk_splats* splats, const float* a, const float* b, uint32_t n_ab,
             const float color[3]);
/* Splats::save_colors / restore_colors (paint.rs:67-93): device copy of the
 * three DC planes (3 x n floats, field-major) to / from caller memory. */
int sfk_save_colors(sfk_ctx* ctx, const sfk_splats* splats, float* saved);
int sfk_restore_colors(sfk_ctx* ctx, sfk_splats* splats, const float* saved);

/* ------------------------------------- batched EIG round (no reference
 * counterpart: the reference scores candidates one plain render at a time,
 * src/seg/active.rs:113-129) */

typedef struct sfk_eig_round sfk_eig_round;

/* A round over nviews (<= 64) candidate cameras of mixed resolution, at one
 * tile width (8, 16, 32) and radix digit width (4..8).  Buffers are sized
 * once for (splats x views). */
int sfk_eig_round_create(sfk_ctx* ctx, const sfk_splats* splats, const sfk_camera* cams,
                         const uint32_t* widths, const uint32_t* heights, uint32_t nviews,
                         uint32_t tile_w, uint32_t digit_bits, sfk_eig_round** out);
void sfk_eig_round_destroy(sfk_eig_round* r);
/* Score every view against the posterior (a, b): responsibility renders,
 * f64 EIG per view and the argmax, as one sequence; use_graph != 0 captures
 * it into a hipGraph on first use (re-captured if a or b move) and replays
 * it.  eig_out (nviews doubles) and best_out, if non-NULL, are filled after
 * one synchronization. */
int sfk_eig_round_run(sfk_ctx* ctx, sfk_eig_round* r, const float* a, const float* b,
                      int use_graph, double* eig_out, uint32_t* best_out);
/* Binned isect emission (default, grids of at most 255 x 255 tiles) or
 * emission + stable tile argsort; the same isect lists either way.  Drops a
 * captured graph. */
int sfk_eig_round_set_binned(sfk_eig_round* r, int on);
/* Binned only: the raster takes each tile's isects straight from its
 * supertile's depth-ordered (splat, bbox) pairs (default), or from per-tile
 * lists placed by the binned emission; the same evidence either way.  No
 * counterpart in the reference (its rounds render views one at a time).
 * Drops a captured graph. */
int sfk_eig_round_set_pair_raster(sfk_eig_round* r, int on);
/* Device pointer of one view's u32 responsibility words and its scale. */
int sfk_eig_round_info(const sfk_eig_round* r, uint32_t view, const uint32_t** resp,
                       float* scale, uint32_t* capacity, uint32_t* total_tiles);
/* SYNC: per view [isects, visible], then the total isect count. */
int sfk_eig_round_counters(sfk_ctx* ctx, const sfk_eig_round* r, uint32_t* counters);
/* Autotuner: time the graph-replayed round at every tile width {8,16,32} x
 * digit width {4,8} (median of `reps` replays, HIP events) and return the
 * fastest; ms_out (6 floats, may be NULL) gets each configuration's time,
 * tile-major.  Evidence depends on the tile width (the bbox cutoff is 10/255,
 * the compositor's 1/255), so compare against plain renders at the same one. */
int sfk_eig_round_autotune(sfk_ctx* ctx, const sfk_splats* splats, const sfk_camera* cams,
                           const uint32_t* widths, const uint32_t* heights, uint32_t nviews,
                           const float* a, const float* b, uint32_t reps, uint32_t* best_tile_w,
                           uint32_t* best_digit_bits, float* ms_out);
/* The plain round (Segmenter::select_view, active.rs:113-129): per view a
 * full render_responsibility and a synchronous sfk_eig, argmax on the host. */
int sfk_eig_round_plain(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                        const sfk_camera* cams, const uint32_t* widths, const uint32_t* heights,
                        uint32_t nviews, const float* a, const float* b, uint32_t* resp,
                        double* eig_out, uint32_t* best_out);

/* ------------------------------------------------ backward (no reference
 * counterpart: the reference renders forward only, docs/00_pipeline.md:1) */

/* L2 image loss: d_rgba = 2 (rgba - target) over pixels x 4 channels;
 * d_loss (device f64, may be NULL) = sum of squared differences. */
int sfk_loss_l2(sfk_ctx* ctx, const float* rgba, const float* target, uint32_t pixels,
                float* d_rgba, double* d_loss);
/* Adjoint of finalize(Rgb) and project_splats for the frame last prepared
 * and composited in `scratch` (its final_t / n_contrib): given dL/d(rgba)
 * (h x w x 4 f32), overwrite grad_attrs (11 planes x n, the attribute
 * layout: means, raw quaternion, log-scales, opacity logit) and grad_sh
 * (3*sh_per_ch planes x n). */
int sfk_render_backward(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                        const sfk_tile_isects* isects, const float* d_rgba, float* grad_attrs,
                        float* grad_sh);
/* sfk_loss_l2 followed by sfk_render_backward in one pass: the adjoint
 * rasterizer forms d_rgba = 2 (rgba - target) per pixel itself (no d_rgba
 * buffer) and writes the L2 loss, a fixed-order f64 sum, to d_loss (device,
 * required).  Gradients equal the two-call sequence bit for bit. */
int sfk_render_backward_l2(sfk_ctx* ctx, const sfk_splats* splats, sfk_render_scratch* scratch,
                           const sfk_tile_isects* isects, const float* rgba, const float* target,
                           float* grad_attrs, float* grad_sh, double* d_loss);

/* ------------------------------------------------------ kernel-level calls */

/* project_splats (src/project.rs:234-327) over all n splats.  Writes per
 * splat i: depth_keys[i], projected row i, tile_counts[i], bbox[2i..2i+1];
 * adds (isects, visible) of each workgroup to one of 64 slot pairs:
 * counters[2s] += isects, counters[2s+1] += visible (128 u32, not cleared;
 * the totals are the sums over s). */
int sfk_project(sfk_ctx* ctx, const sfk_view* view, const float* attrs, const float* sh,
                uint32_t n, uint32_t sh_per_ch, uint32_t tile_w, uint32_t* depth_keys,
                float* projected, uint32_t* tile_counts, uint32_t* bbox, uint32_t* counters);
/* map_isects (src/raster.rs:12-48) over ranks [0, n) of depth_order. */
int sfk_map_isects(sfk_ctx* ctx, const uint32_t* depth_order, const uint32_t* tile_counts,
                   const uint32_t* bbox, const uint32_t* offsets, uint32_t n,
                   uint32_t tiles_per_row, uint32_t max_isects, uint32_t* tile_ids,
                   uint32_t* gaussian_ids);
/* lower_bound tile starts (src/raster.rs:52-65, 111-115) for all tiles in
 * [0, num_tiles]; n isects, or min(*d_n, n) when d_n is given. */
int sfk_tile_starts(sfk_ctx* ctx, const uint32_t* tile_ids, uint32_t n, const uint32_t* d_n,
                    uint32_t num_tiles, uint32_t* starts);

#ifdef __cplusplus
}
#endif
#endif
