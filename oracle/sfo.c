/* Oracle (test infrastructure only): CPU restatement of the reference's 3DGS
 * forward render and B3-Seg evidence path.  See sfo.h for the rules of use.
 *
 * Arithmetic follows the reference operation by operation (left-to-right
 * evaluation, no contraction: this file is built with -ffp-contract=off).
 * Where the reference leaves a choice to the device (transcendentals, the
 * racy compaction order) the choices are fixed here:
 *   - exp/ln: Cephes polynomials (sfo_math.h); sqrt and division: IEEE
 *     correctly rounded; inverseSqrt(x) -> 1/sqrt(x); recip(x) -> 1/x.
 *   - visible splats are ranked by (depth key, splat index): the stable depth
 *     sort of the reference fed with index-ordered compaction, one of the
 *     orders the reference itself can produce (src/project.rs:50-54, 281-283).
 */
#include "sfo.h"
#include "sfo_math.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define SFO_PROJ_FLOATS 9 /* src/layout.rs:60-61 */

/* src/project.rs:13 — 10/255 evaluated in f32. */
static const float SFO_ALPHA_CUTOFF = 10.0f / 255.0f;

/* ---------------------------------------------------------------- camera */

/* Rotation matrix columns of a unit quaternion (glam Mat3::from_quat form,
 * as used by src/camera.rs:42-44). */
static void quat_cols(const float q[4], float c0[3], float c1[3], float c2[3]) {
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float x2 = x + x, y2 = y + y, z2 = z + z;
    float xx = x * x2, xy = x * y2, xz = x * z2;
    float yy = y * y2, yz = y * z2, zz = z * z2;
    float wx = w * x2, wy = w * y2, wz = w * z2;
    c0[0] = 1.0f - (yy + zz); c0[1] = xy + wz;            c0[2] = xz - wy;
    c1[0] = xy - wz;          c1[1] = 1.0f - (xx + zz);   c1[2] = yz + wx;
    c2[0] = xz + wy;          c2[1] = yz - wx;            c2[2] = 1.0f - (xx + yy);
}

static float dot3(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* CameraView::for_camera (src/project.rs:76-96) with w2c from
 * src/camera.rs:42-44 taken as the exact rigid inverse (R^T, -R^T p). */
void sfo_camera_view(const float fov[2], const float pos[3], const float q[4], uint32_t w,
                     uint32_t h, uint32_t tile_w, sfo_view* out) {
    quat_cols(q, out->rot0, out->rot1, out->rot2);
    out->trans[0] = -dot3(out->rot0, pos);
    out->trans[1] = -dot3(out->rot1, pos);
    out->trans[2] = -dot3(out->rot2, pos);
    /* src/camera.rs:38-40 */
    out->focal[0] = ((float)w * 0.5f) / tanf(fov[0] * 0.5f);
    out->focal[1] = ((float)h * 0.5f) / tanf(fov[1] * 0.5f);
    out->world_pos[0] = pos[0];
    out->world_pos[1] = pos[1];
    out->world_pos[2] = pos[2];
    out->img[0] = (float)w;
    out->img[1] = (float)h;
    /* src/render.rs:137-141 */
    out->tile_bounds[0] = (float)((w + tile_w - 1) / tile_w);
    out->tile_bounds[1] = (float)((h + tile_w - 1) / tile_w);
}

/* Camera::look_at (src/camera.rs:57-63): minimal arc from +Z to the view
 * direction; the antiparallel case turns pi about +Y (glam's choice of the
 * orthonormal axis for +Z). */
void sfo_look_at(const float pos[3], const float target[3], float q[4]) {
    float d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    float inv = 1.0f / sqrtf(dot3(d, d));
    float dir[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
    const float eps = 1.1920929e-7f;
    float dp = dir[2];
    if (dp > 1.0f - 2.0f * eps) {
        q[0] = 0.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 1.0f;
    } else if (dp < -1.0f + 2.0f * eps) {
        float half = 3.14159265358979323846f * 0.5f;
        q[0] = 0.0f; q[1] = sinf(half); q[2] = 0.0f; q[3] = cosf(half);
    } else {
        float c[4] = {-dir[1], dir[0], 0.0f, 1.0f + dp};
        float len = sqrtf((c[0] * c[0] + c[1] * c[1]) + (c[2] * c[2] + c[3] * c[3]));
        float r = 1.0f / len;
        q[0] = c[0] * r; q[1] = c[1] * r; q[2] = c[2] * r; q[3] = c[3] * r;
    }
}

/* crates/sort/src/sort.rs:259-261 */
uint32_t sfo_bits_for(uint32_t max_exclusive) {
    uint32_t v = max_exclusive == 0 ? 0 : max_exclusive - 1;
    uint32_t b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

/* src/seg/mod.rs:31-34 */
uint32_t sfo_fractional_bits(uint32_t w, uint32_t h) {
    uint64_t p = (uint64_t)w * (uint64_t)h;
    uint64_t np2 = 1;
    uint32_t tz = 0;
    while (np2 < p) { np2 <<= 1; tz++; }
    return tz >= 32 ? 0 : 32 - tz;
}

/* --------------------------------------------------------------- project */

static float sigmoidf_(float x) { return 1.0f / (1.0f + sfo_expf(-x)); } /* project.rs:32-35 */

#define SH_C0 0.2820948f
#define SH_C1 0.48860252f
#define SH_C2_0 0.9461747f
#define SH_C2_1 0.31539157f
#define SH_C2_2 -1.0925485f
#define SH_C2_3 0.54627424f

/* src/project.rs:342-388, including its band-3 basis as written. */
static void sh_to_rgb(uint32_t chs, const float dir[3], uint32_t s, uint32_t n,
                      const float* shs, float rgb[3]) {
    size_t st = n;
#define SH(k) shs[(size_t)(k) * st + s]
    float x = dir[0], y = dir[1], z = dir[2];
    float r = SH_C0 * SH(0);
    float g = SH_C0 * SH(1);
    float b = SH_C0 * SH(2);
    if (chs >= 4) {
        r += SH_C1 * ((-y * SH(3) + z * SH(6)) - x * SH(9));
        g += SH_C1 * ((-y * SH(4) + z * SH(7)) - x * SH(10));
        b += SH_C1 * ((-y * SH(5) + z * SH(8)) - x * SH(11));
    }
    if (chs >= 9) {
        float b4 = SH_C2_2 * z * x;
        float b5 = SH_C2_2 * z * y;
        float b6 = SH_C2_0 * z * z - SH_C2_1;
        float b7 = SH_C2_3 * 2.0f * x * y;
        float b8 = SH_C2_3 * (x * x - y * y);
        r += b6 * SH(18) + b7 * SH(12) + b5 * SH(15) + b4 * SH(21) + b8 * SH(24);
        g += b6 * SH(19) + b7 * SH(13) + b5 * SH(16) + b4 * SH(22) + b8 * SH(25);
        b += b6 * SH(20) + b7 * SH(14) + b5 * SH(17) + b4 * SH(23) + b8 * SH(26);
    }
    if (chs >= 16) {
        float c1x = x * x - y * y;
        float c1y = 2.0f * x * y;
        float tmp0c = -2.285229f * z * z + 0.4570458f;
        float tmp1b = 1.4453057f * z;
        float b9 = -0.5900436f * x * c1y + y * c1x;
        float b10 = -0.5900436f * x * c1x - y * c1y;
        float b11 = tmp0c * y;
        float b12 = tmp0c * x;
        float b13 = z * (1.8658817f * z * z - 1.119529f);
        float b14 = tmp1b * c1y;
        float b15 = tmp1b * c1x;
        r += b9 * SH(27) + b10 * SH(30) + b11 * SH(33) + b12 * SH(36) + b13 * SH(39) +
             b14 * SH(42) + b15 * SH(45);
        g += b9 * SH(28) + b10 * SH(31) + b11 * SH(34) + b12 * SH(37) + b13 * SH(40) +
             b14 * SH(43) + b15 * SH(46);
        b += b9 * SH(29) + b10 * SH(32) + b11 * SH(35) + b12 * SH(38) + b13 * SH(41) +
             b14 * SH(44) + b15 * SH(47);
    }
#undef SH
    rgb[0] = r; rgb[1] = g; rgb[2] = b;
}

typedef struct {
    const sfo_view* v;
    const float* attrs;
    const float* sh;
    uint32_t n, sh_per_ch, tile_w, lo, hi;
    uint8_t* visible;
    uint32_t* depth_keys;
    float* projected;
    uint32_t* tile_counts;
    uint32_t* bbox;
    uint64_t isects, vis;
} project_job;

/* One splat of project_splats (src/project.rs:234-388). */
static void project_one(project_job* J, uint32_t i) {
    const sfo_view* v = J->v;
    size_t n = J->n;
    const float* A = J->attrs;
    float mean[3] = {A[0 * n + i], A[1 * n + i], A[2 * n + i]};
    float qw = A[3 * n + i], qx = A[4 * n + i], qy = A[5 * n + i], qz = A[6 * n + i];
    float sc[3] = {sfo_expf(A[7 * n + i]), sfo_expf(A[8 * n + i]), sfo_expf(A[9 * n + i])};
    float opacity = sigmoidf_(A[10 * n + i]);
    J->visible[i] = 0;
    J->tile_counts[i] = 0;
    if (opacity < SFO_ALPHA_CUTOFF) return;

    float cam[3];
    cam[0] = dot3(v->rot0, mean) + v->trans[0];
    cam[1] = dot3(v->rot1, mean) + v->trans[1];
    cam[2] = dot3(v->rot2, mean) + v->trans[2];
    if (cam[2] <= 0.1f) return;
    float inv_z = 1.0f / cam[2];

    /* compute_cov2d (project.rs:175-221) */
    float x2 = qx * qx, y2 = qy * qy, z2 = qz * qz;
    float xy = qx * qy, xz = qx * qz, yz = qy * qz;
    float wx = qw * qx, wy = qw * qy, wz = qw * qz;
    float r0[3] = {1.0f - 2.0f * (y2 + z2), 2.0f * (xy - wz), 2.0f * (xz + wy)};
    float r1[3] = {2.0f * (xy + wz), 1.0f - 2.0f * (x2 + z2), 2.0f * (yz - wx)};
    float r2[3] = {2.0f * (xz - wy), 2.0f * (yz + wx), 1.0f - 2.0f * (x2 + y2)};
    float m0[3] = {r0[0] * sc[0], r0[1] * sc[1], r0[2] * sc[2]};
    float m1[3] = {r1[0] * sc[0], r1[1] * sc[1], r1[2] * sc[2]};
    float m2[3] = {r2[0] * sc[0], r2[1] * sc[1], r2[2] * sc[2]};
    float lim_x = 1.3f * v->img[0] / (2.0f * v->focal[0]);
    float lim_y = 1.3f * v->img[1] / (2.0f * v->focal[1]);
    float u = sfo_clampf(cam[0] * inv_z, -lim_x, lim_x);
    float vv = sfo_clampf(cam[1] * inv_z, -lim_y, lim_y);
    float fx_iz = v->focal[0] * inv_z;
    float fu_iz = fx_iz * u;
    float t0[3] = {fx_iz * v->rot0[0] - fu_iz * v->rot2[0], fx_iz * v->rot0[1] - fu_iz * v->rot2[1],
                   fx_iz * v->rot0[2] - fu_iz * v->rot2[2]};
    float fy_iz = v->focal[1] * inv_z;
    float fv_iz = fy_iz * vv;
    float t1[3] = {fy_iz * v->rot1[0] - fv_iz * v->rot2[0], fy_iz * v->rot1[1] - fv_iz * v->rot2[1],
                   fy_iz * v->rot1[2] - fv_iz * v->rot2[2]};
    float j0[3], j1[3];
    for (int k = 0; k < 3; k++) {
        j0[k] = t0[0] * m0[k] + t0[1] * m1[k] + t0[2] * m2[k];
        j1[k] = t1[0] * m0[k] + t1[1] * m1[k] + t1[2] * m2[k];
    }
    float ca = dot3(j0, j0) + 0.3f;
    float cb = dot3(j0, j1);
    float cc = dot3(j1, j1) + 0.3f;
    /* compute_conic (project.rs:223-232) */
    float det = fmaxf(ca * cc - cb * cb, 1e-6f);
    float inv_det = 1.0f / det;

    J->visible[i] = 1;
    J->depth_keys[i] = sfo_f2b(cam[2]); /* project.rs:56-58, 283 */

    float dir[3] = {mean[0] - v->world_pos[0], mean[1] - v->world_pos[1], mean[2] - v->world_pos[2]};
    float inv = 1.0f / sqrtf(dot3(dir, dir));
    float nd[3] = {dir[0] * inv, dir[1] * inv, dir[2] * inv};
    float rgb[3];
    sh_to_rgb(J->sh_per_ch, nd, i, J->n, J->sh, rgb);

    float mx = v->focal[0] * cam[0] * inv_z + v->img[0] * 0.5f;
    float my = v->focal[1] * cam[1] * inv_z + v->img[1] * 0.5f;
    float* P = J->projected + (size_t)i * SFO_PROJ_FLOATS;
    P[0] = mx;
    P[1] = my;
    P[2] = cc * inv_det;
    P[3] = -cb * inv_det;
    P[4] = ca * inv_det;
    P[5] = rgb[0] + 0.5f;
    P[6] = rgb[1] + 0.5f;
    P[7] = rgb[2] + 0.5f;
    P[8] = opacity;

    float cutoff = sfo_logf(opacity / SFO_ALPHA_CUTOFF);
    float ex = sqrtf(2.0f * cutoff * ca);
    float ey = sqrtf(2.0f * cutoff * cc);
    /* tile_bbox (project.rs:37-46) */
    float inv_tile = 1.0f / (float)J->tile_w;
    float bx = v->tile_bounds[0], by = v->tile_bounds[1];
    uint32_t min_x = sfo_f2u(sfo_clampf((mx - ex) * inv_tile, 0.0f, bx));
    uint32_t min_y = sfo_f2u(sfo_clampf((my - ey) * inv_tile, 0.0f, by));
    uint32_t max_x = sfo_f2u(sfo_clampf((mx + ex) * inv_tile + 1.0f, 0.0f, bx));
    uint32_t max_y = sfo_f2u(sfo_clampf((my + ey) * inv_tile + 1.0f, 0.0f, by));
    uint32_t cnt = (max_x - min_x) * (max_y - min_y);
    J->tile_counts[i] = cnt;
    J->isects += cnt;
    J->vis += 1;
    J->bbox[2 * (size_t)i] = min_x | (min_y << 16);
    J->bbox[2 * (size_t)i + 1] = max_x | (max_y << 16);
}

static void* project_worker(void* arg) {
    project_job* J = (project_job*)arg;
    for (uint32_t i = J->lo; i < J->hi; i++) project_one(J, i);
    return NULL;
}

void sfo_project(const sfo_view* v, const float* attrs, const float* sh, uint32_t n,
                 uint32_t sh_per_ch, uint32_t tile_w, int nthreads, uint8_t* visible,
                 uint32_t* depth_keys, float* projected, uint32_t* tile_counts, uint32_t* bbox,
                 uint32_t counters[2]) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    project_job jobs[256];
    pthread_t th[256];
    uint32_t per = (n + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        project_job* J = &jobs[t];
        J->v = v; J->attrs = attrs; J->sh = sh; J->n = n; J->sh_per_ch = sh_per_ch;
        J->tile_w = tile_w;
        J->lo = (uint32_t)t * per < n ? (uint32_t)t * per : n;
        J->hi = J->lo + per < n ? J->lo + per : n;
        J->visible = visible; J->depth_keys = depth_keys; J->projected = projected;
        J->tile_counts = tile_counts; J->bbox = bbox; J->isects = 0; J->vis = 0;
        if (nthreads > 1) pthread_create(&th[t], NULL, project_worker, J);
        else project_worker(J);
    }
    uint64_t is = 0, vs = 0;
    for (int t = 0; t < nthreads; t++) {
        if (nthreads > 1) pthread_join(th[t], NULL);
        is += jobs[t].isects;
        vs += jobs[t].vis;
    }
    counters[0] = (uint32_t)is; /* u32 atomic counter, wraps like the reference */
    counters[1] = (uint32_t)vs;
}

/* ------------------------------------------------------------------ sort */

/* Stable LSD argsort on the low `bits` of each key (crates/sort/src/sort.rs:
 * 302-381 semantics: equal keys keep input order).  8-bit counting passes. */
void sfo_argsort_u32(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t bits,
                     uint32_t* out_keys, uint32_t* out_vals) {
    memcpy(out_keys, keys, (size_t)n * 4);
    memcpy(out_vals, vals, (size_t)n * 4);
    if (n <= 1 || bits == 0) return;
    uint32_t* tk = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* tv = (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        uint32_t w = bits - shift < 8 ? bits - shift : 8;
        uint32_t mask = (1u << w) - 1u;
        size_t cnt[257] = {0};
        for (uint32_t i = 0; i < n; i++) cnt[((out_keys[i] >> shift) & mask) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; i++) {
            size_t p = cnt[(out_keys[i] >> shift) & mask]++;
            tk[p] = out_keys[i];
            tv[p] = out_vals[i];
        }
        memcpy(out_keys, tk, (size_t)n * 4);
        memcpy(out_vals, tv, (size_t)n * 4);
    }
    free(tk);
    free(tv);
}

void sfo_argsort_u64(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t bits,
                     uint64_t* out_keys, uint32_t* out_vals) {
    memcpy(out_keys, keys, (size_t)n * 8);
    memcpy(out_vals, vals, (size_t)n * 4);
    if (n <= 1 || bits == 0) return;
    uint64_t* tk = (uint64_t*)malloc((size_t)n * 8);
    uint32_t* tv = (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        uint32_t w = bits - shift < 8 ? bits - shift : 8;
        uint64_t mask = (1ull << w) - 1ull;
        size_t cnt[257] = {0};
        for (uint32_t i = 0; i < n; i++) cnt[((out_keys[i] >> shift) & mask) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; i++) {
            size_t p = cnt[(out_keys[i] >> shift) & mask]++;
            tk[p] = out_keys[i];
            tv[p] = out_vals[i];
        }
        memcpy(out_keys, tk, (size_t)n * 8);
        memcpy(out_vals, tv, (size_t)n * 4);
    }
    free(tk);
    free(tv);
}

/* Visible splats in index order, stable-sorted by depth key
 * (src/render.rs:308-316 with deterministic compaction).  Returns V. */
uint32_t sfo_depth_order(const uint8_t* visible, const uint32_t* keys, uint32_t n,
                         uint32_t* depth_order) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; i++) v += visible[i] ? 1 : 0;
    uint32_t* k = (uint32_t*)malloc((size_t)(v ? v : 1) * 4);
    uint32_t* ids = (uint32_t*)malloc((size_t)(v ? v : 1) * 4);
    uint32_t* ok = (uint32_t*)malloc((size_t)(v ? v : 1) * 4);
    uint32_t j = 0;
    for (uint32_t i = 0; i < n; i++)
        if (visible[i]) { k[j] = keys[i]; ids[j] = i; j++; }
    sfo_argsort_u32(k, ids, v, 32, ok, depth_order);
    free(k); free(ids); free(ok);
    return v;
}

/* ------------------------------------------------------------------ scan */

/* exclusive_scan_gather (crates/sort/src/scan.rs:215-255). */
void sfo_scan_gather(const uint32_t* gids, const uint32_t* counts, uint32_t n,
                     uint32_t* offsets) {
    uint32_t run = 0;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = run;
        run += counts[gids[i]];
    }
}

/* ------------------------------------------------------------------- map */

/* map_isects (src/raster.rs:12-48). */
void sfo_map_isects(const uint32_t* depth_order, const uint32_t* bbox, const uint32_t* offsets,
                    uint32_t tiles_per_row, uint32_t max_isects, uint32_t num_visible,
                    uint32_t* tile_ids, uint32_t* gaussian_ids) {
    for (uint32_t r = 0; r < num_visible; r++) {
        uint32_t vis = depth_order[r];
        uint32_t lo = bbox[2 * (size_t)vis], hi = bbox[2 * (size_t)vis + 1];
        uint32_t min_x = lo & 0xffffu, min_y = lo >> 16, max_x = hi & 0xffffu, max_y = hi >> 16;
        uint32_t slot = offsets[r];
        for (uint32_t ty = min_y; ty < max_y; ty++)
            for (uint32_t tx = min_x; tx < max_x; tx++) {
                if (slot < max_isects) {
                    tile_ids[slot] = tx + ty * tiles_per_row;
                    gaussian_ids[slot] = vis;
                }
                slot++;
            }
    }
}

/* lower_bound (src/raster.rs:52-65). */
uint32_t sfo_lower_bound(const uint32_t* ids, uint32_t n, uint32_t target) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (uint32_t)(((uint64_t)lo + hi) / 2);
        if (ids[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* Per-tile [start, end) of the sorted run (src/raster.rs:111-115). */
void sfo_tile_ranges(const uint32_t* tile_ids, uint32_t num_isects, uint32_t num_tiles,
                     uint32_t* ranges) {
    for (uint32_t t = 0; t < num_tiles; t++) {
        ranges[2 * (size_t)t] = sfo_lower_bound(tile_ids, num_isects, t);
        ranges[2 * (size_t)t + 1] = sfo_lower_bound(tile_ids, num_isects, t + 1);
    }
}

/* ---------------------------------------------------------------- raster */

typedef struct {
    int mode;
    uint32_t w, h, tile_w, row_stride, tiles_x, ty_lo, ty_hi;
    float scale;
    const uint32_t *ranges, *gids, *mask;
    const float* proj;
    uint32_t words_per_row;
    float* rgba;
    uint32_t* bitmap;
    float* final_t;
    uint32_t* n_contrib;
    uint32_t *resp, *fg, *bg;
} raster_job;

static uint32_t quantize_u8(float v) { return sfo_f2u(sfo_clampf(v * 255.0f, 0.0f, 255.0f)); }

/* rasterize_kernel, one pixel (src/raster.rs:87-230). */
static void raster_pixel(raster_job* J, uint32_t px, uint32_t py, uint32_t start, uint32_t end) {
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
    uint32_t inside = 0;
    if (J->mode == SFO_MODE_EVIDENCE)
        inside = (J->mask[px / 32 + py * J->words_per_row] >> (px % 32)) & 1u;
    uint32_t last = start;
    for (uint32_t k = start; k < end; k++) {
        uint32_t g = J->gids[k];
        const float* P = J->proj + (size_t)g * SFO_PROJ_FLOATS;
        float dx = P[0] - pxf, dy = P[1] - pyf;
        float power = 0.5f * (P[2] * dx * dx + P[4] * dy * dy) + P[3] * dx * dy;
        float alpha = fminf(P[8] * sfo_expf(-power), 0.999f);
        if (alpha >= 1.0f / 255.0f) {
            float vis = alpha * T;
            if (J->mode == SFO_MODE_RGB) {
                cr += P[5] * vis;
                cg += P[6] * vis;
                cb += P[7] * vis;
            } else {
                uint32_t q = sfo_f2u(vis * J->scale);
                uint32_t* dst = J->mode == SFO_MODE_TOTAL ? J->resp : (inside ? J->fg : J->bg);
                __atomic_fetch_add(&dst[g], q, __ATOMIC_RELAXED);
            }
            T *= 1.0f - alpha;
            last = k + 1;
            if (T < 1.0f / 255.0f) break;
        }
    }
    size_t p = (size_t)py * J->w + px;
    if (J->final_t) J->final_t[p] = T;
    if (J->n_contrib) J->n_contrib[p] = last;
    if (J->mode == SFO_MODE_RGB) {
        if (J->rgba) {
            J->rgba[4 * p + 0] = cr;
            J->rgba[4 * p + 1] = cg;
            J->rgba[4 * p + 2] = cb;
            J->rgba[4 * p + 3] = 1.0f - T;
        }
        if (J->bitmap) {
            uint32_t r = quantize_u8(cr), g = quantize_u8(cg), b = quantize_u8(cb),
                     a = quantize_u8(1.0f - T);
            J->bitmap[px + (size_t)py * J->row_stride] = r | (g << 8) | (b << 16) | (a << 24);
        }
    }
}

static void* raster_worker(void* arg) {
    raster_job* J = (raster_job*)arg;
    for (uint32_t ty = J->ty_lo; ty < J->ty_hi; ty++)
        for (uint32_t tx = 0; tx < J->tiles_x; tx++) {
            uint32_t tile = tx + ty * J->tiles_x;
            uint32_t start = J->ranges[2 * (size_t)tile], end = J->ranges[2 * (size_t)tile + 1];
            for (uint32_t yy = 0; yy < J->tile_w; yy++)
                for (uint32_t xx = 0; xx < J->tile_w; xx++) {
                    uint32_t px = tx * J->tile_w + xx, py = ty * J->tile_w + yy;
                    if (px < J->w && py < J->h) raster_pixel(J, px, py, start, end);
                }
        }
    return NULL;
}

void sfo_rasterize(int mode, uint32_t w, uint32_t h, uint32_t tile_w, uint32_t row_stride,
                   float scale, const uint32_t* ranges, const uint32_t* gaussian_ids,
                   const float* projected, const uint32_t* mask, uint32_t words_per_row,
                   int nthreads, float* rgba, uint32_t* bitmap, float* final_t,
                   uint32_t* n_contrib, uint32_t* resp, uint32_t* fg, uint32_t* bg) {
    uint32_t tiles_x = (w + tile_w - 1) / tile_w, tiles_y = (h + tile_w - 1) / tile_w;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > tiles_y) nthreads = (int)tiles_y;
    raster_job jobs[256];
    pthread_t th[256];
    uint32_t per = (tiles_y + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        raster_job* J = &jobs[t];
        J->mode = mode; J->w = w; J->h = h; J->tile_w = tile_w; J->row_stride = row_stride;
        J->tiles_x = tiles_x;
        J->ty_lo = (uint32_t)t * per < tiles_y ? (uint32_t)t * per : tiles_y;
        J->ty_hi = J->ty_lo + per < tiles_y ? J->ty_lo + per : tiles_y;
        J->scale = scale; J->ranges = ranges; J->gids = gaussian_ids; J->mask = mask;
        J->proj = projected; J->words_per_row = words_per_row; J->rgba = rgba;
        J->bitmap = bitmap; J->final_t = final_t; J->n_contrib = n_contrib;
        J->resp = resp; J->fg = fg; J->bg = bg;
        if (nthreads > 1) pthread_create(&th[t], NULL, raster_worker, J);
        else raster_worker(J);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------- seg */

/* ln Gamma, Lanczos g = 7 with 9 terms (src/seg/beta.rs:52-72). */
double sfo_lgamma(double x) {
    static const double c[8] = {676.5203681218851,     -1259.1392167224028,
                                771.3234287776531,     -176.6150291621406,
                                12.507343278686905,    -0.13857109526572012,
                                9.984369578019572e-6,  1.5056327351493116e-7};
    double z = x - 1.0;
    double series = 0.9999999999998099;
    for (int k = 0; k < 8; k++) series += c[k] / (z + (double)k + 1.0);
    double t = z + 7.5;
    return 0.5 * log(2.0 * 3.14159265358979323846) + (z + 0.5) * log(t) - t + log(series);
}

/* digamma: recurrence to x >= 6, then the asymptotic series
 * (src/seg/beta.rs:74-86). */
double sfo_digamma(double x) {
    double r = 0.0;
    while (x < 6.0) {
        r -= 1.0 / x;
        x += 1.0;
    }
    double inv = 1.0 / x, inv2 = inv * inv;
    return r + log(x) - 0.5 * inv - inv2 / 12.0 + inv2 * inv2 / 120.0 -
           inv2 * inv2 * inv2 / 252.0;
}

/* Differential entropy of Beta(a, b) (src/seg/beta.rs:88-96). */
double sfo_beta_entropy(double a, double b) {
    double s = a + b;
    return sfo_lgamma(a) + sfo_lgamma(b) - sfo_lgamma(s) - (a - 1.0) * sfo_digamma(a) -
           (b - 1.0) * sfo_digamma(b) + (s - 2.0) * sfo_digamma(s);
}

/* EIG of one candidate: sum_i max(0, H(a,b) - H(a k, b k)), k = 1 + eps/(a+b),
 * eps = f32(q)/scale, skipped at eps = 0 (src/seg/beta.rs:146-201, 240-268),
 * with H evaluated in f64 instead of the f32 table. */
double sfo_eig(const float* a, const float* b, const uint32_t* resp, uint32_t n, float scale) {
    double total = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        float eps = (float)resp[i] / scale;
        if (eps > 0.0f) {
            double ai = a[i], bi = b[i];
            double k = 1.0 + (double)eps / (ai + bi);
            double d = sfo_beta_entropy(ai, bi) - sfo_beta_entropy(ai * k, bi * k);
            total += d > 0.0 ? d : 0.0;
        }
    }
    return total;
}

/* beta_update_kernel (src/seg/beta.rs:203-212). */
void sfo_beta_update(float scale, const uint32_t* fg, const uint32_t* bg, uint32_t n, float* a,
                     float* b) {
    for (uint32_t i = 0; i < n; i++) {
        a[i] += (float)fg[i] / scale;
        b[i] += (float)bg[i] / scale;
    }
}

/* paint_posterior (src/seg/paint.rs:99-109): row colour := a/(a+b) at the
 * reference row's offsets 5..7, no +0.5. */
void sfo_paint_posterior(float* projected, const float* a, const float* b, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        float m = a[i] / (a[i] + b[i]);
        float* P = projected + (size_t)i * SFO_PROJ_FLOATS;
        P[5] = m;
        P[6] = m;
        P[7] = m;
    }
}

/* tint_colors (src/seg/paint.rs:113-122): DC channel c of splat i at
 * c*n + i, overwritten where a > b. */
void sfo_tint_colors(float* sh, const float* a, const float* b, uint32_t n, float r, float g,
                     float bl) {
    for (uint32_t i = 0; i < n; i++)
        if (a[i] > b[i]) {
            sh[i] = r;
            sh[(size_t)n + i] = g;
            sh[2 * (size_t)n + i] = bl;
        }
}

float sfo_expf_export(float x) { return sfo_expf(x); }
float sfo_logf_export(float x) { return sfo_logf(x); }
