/* Oracle (test infrastructure only — imported by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, never by the product): a plain-C
 * restatement of the reference's 3DGS render path and B3-Seg evidence math.
 *
 * Every function cites the reference file:line it follows.  The product
 * (splatfield_amd/, libsfk.so) must match these results bit for bit on
 * integer outputs (sort order, tile ranges, overlap counts, u32 evidence) and
 * on the f32 values it shares with them.
 */
#ifndef SFO_H
#define SFO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel view of a camera (src/project.rs:60-74, CameraView). */
typedef struct {
    float rot0[3], rot1[3], rot2[3]; /* rows of the world->camera rotation */
    float trans[3];                  /* world->camera translation */
    float focal[2];
    float world_pos[3];
    float img[2];
    float tile_bounds[2];
} sfo_view;

enum { SFO_MODE_RGB = 0, SFO_MODE_TOTAL = 1, SFO_MODE_EVIDENCE = 2 };

void sfo_camera_view(const float fov[2], const float pos[3], const float quat_xyzw[4],
                     uint32_t w, uint32_t h, uint32_t tile_w, sfo_view* out);
void sfo_look_at(const float pos[3], const float target[3], float quat_xyzw_out[4]);

uint32_t sfo_bits_for(uint32_t max_exclusive);
uint32_t sfo_fractional_bits(uint32_t w, uint32_t h);

void sfo_project(const sfo_view* v, const float* attrs, const float* sh, uint32_t n,
                 uint32_t sh_per_ch, uint32_t tile_w, int nthreads, uint8_t* visible,
                 uint32_t* depth_keys, float* projected, uint32_t* tile_counts,
                 uint32_t* bbox, uint32_t counters[2]);

uint32_t sfo_depth_order(const uint8_t* visible, const uint32_t* keys, uint32_t n,
                         uint32_t* depth_order);

void sfo_argsort_u32(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t bits,
                     uint32_t* out_keys, uint32_t* out_vals);
void sfo_argsort_u64(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t bits,
                     uint64_t* out_keys, uint32_t* out_vals);

void sfo_scan_gather(const uint32_t* gids, const uint32_t* counts, uint32_t n,
                     uint32_t* offsets);

void sfo_map_isects(const uint32_t* depth_order, const uint32_t* bbox, const uint32_t* offsets,
                    uint32_t tiles_per_row, uint32_t max_isects, uint32_t num_visible,
                    uint32_t* tile_ids, uint32_t* gaussian_ids);

uint32_t sfo_lower_bound(const uint32_t* ids, uint32_t n, uint32_t target);
void sfo_tile_ranges(const uint32_t* tile_ids, uint32_t num_isects, uint32_t num_tiles,
                     uint32_t* ranges);

void sfo_rasterize(int mode, uint32_t w, uint32_t h, uint32_t tile_w, uint32_t row_stride,
                   float scale, const uint32_t* ranges, const uint32_t* gaussian_ids,
                   const float* projected, const uint32_t* mask, uint32_t words_per_row,
                   int nthreads, float* rgba, uint32_t* bitmap, float* final_t,
                   uint32_t* n_contrib, uint32_t* resp, uint32_t* fg, uint32_t* bg);

double sfo_lgamma(double x);
double sfo_digamma(double x);
double sfo_beta_entropy(double a, double b);
double sfo_eig(const float* a, const float* b, const uint32_t* resp, uint32_t n, float scale);
void sfo_paint_posterior(float* projected, const float* a, const float* b, uint32_t n);
void sfo_tint_colors(float* sh, const float* a, const float* b, uint32_t n, float r, float g,
                     float bl);
void sfo_beta_update(float scale, const uint32_t* fg, const uint32_t* bg, uint32_t n,
                     float* a, float* b);

float sfo_expf_export(float x);
float sfo_logf_export(float x);

#ifdef __cplusplus
}
#endif
#endif
