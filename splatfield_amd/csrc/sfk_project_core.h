// Per-splat projection math of project_splats (src/project.rs:234-388),
// shared by the single-view and the batched multi-view kernels so that both
// produce the same bits.  Operation order follows the reference; the library
// is compiled with -ffp-contract=off.
#pragma once

#include "sfk_device.h"
#include "sfk_internal.h"
#include "sfk_math.h"

namespace sfk {

constexpr float kAlphaCutoff = 10.0f / 255.0f;  // project.rs:13

struct SplatIn {
    float mx, my, mz, qw, qx, qy, qz, sx, sy, sz, opacity;
};

struct ProjGeom {
    float px, py;          // mean2d
    float qa, qb, qc;      // conic
    float cz;              // camera depth
    uint32_t bb_lo, bb_hi; // packed tile bbox
    uint32_t count;        // tiles covered
};

__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// Loads splat i of the field-major planes, applying exp/sigmoid as the
// reference does before its culls (project.rs:245-271).
// The 11 loads are issued together and their values pinned here: the
// scales are only used past the culls, and left to itself the compiler sinks
// their loads below the cull branches, one memory round trip each.
__device__ __forceinline__ SplatIn load_splat(const float* __restrict__ A, uint32_t i, uint32_t n) {
    const size_t N = n;
    float r[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) r[k] = A[(size_t)k * N + i];
#pragma unroll
    for (int k = 0; k < 11; ++k) asm volatile("" : "+v"(r[k]));
    SplatIn s;
    s.mx = r[0]; s.my = r[1]; s.mz = r[2];
    s.qw = r[3]; s.qx = r[4]; s.qy = r[5]; s.qz = r[6];
    s.sx = sfk_expf(r[7]);
    s.sy = sfk_expf(r[8]);
    s.sz = sfk_expf(r[9]);
    s.opacity = 1.0f / (1.0f + sfk_expf(-r[10]));
    return s;
}

// The two culls of project_geom alone (opacity, then camera depth).
__device__ __forceinline__ bool passes_culls(const sfk_view& v, const SplatIn& s) {
    if (s.opacity < kAlphaCutoff) return false;
    const float cz = dot3f(v.rot2[0], v.rot2[1], v.rot2[2], s.mx, s.my, s.mz) + v.trans[2];
    return !(cz <= 0.1f);
}

// The view-independent part of compute_cov2d (project.rs:175-221): the
// rotation of the quaternion times the scales, M = R S, row-major.
__device__ __forceinline__ void splat_rs(const SplatIn& s, float (&m)[9]) {
    float x2 = s.qx * s.qx, y2 = s.qy * s.qy, z2 = s.qz * s.qz;
    float xy = s.qx * s.qy, xz = s.qx * s.qz, yz = s.qy * s.qz;
    float wx = s.qw * s.qx, wy = s.qw * s.qy, wz = s.qw * s.qz;
    float r00 = 1.0f - 2.0f * (y2 + z2), r01 = 2.0f * (xy - wz), r02 = 2.0f * (xz + wy);
    float r10 = 2.0f * (xy + wz), r11 = 1.0f - 2.0f * (x2 + z2), r12 = 2.0f * (yz - wx);
    float r20 = 2.0f * (xz - wy), r21 = 2.0f * (yz + wx), r22 = 1.0f - 2.0f * (x2 + y2);
    m[0] = r00 * s.sx; m[1] = r01 * s.sy; m[2] = r02 * s.sz;
    m[3] = r10 * s.sx; m[4] = r11 * s.sy; m[5] = r12 * s.sz;
    m[6] = r20 * s.sx; m[7] = r21 * s.sy; m[8] = r22 * s.sz;
}

// The extent cutoff ln(opacity / (1/255)) of project.rs (view-independent).
__device__ __forceinline__ float splat_cutoff(const SplatIn& s) {
    return sfk_logf(s.opacity / kAlphaCutoff);
}

// The clamp limits of the projected mean's tangent, per view.
__device__ __forceinline__ float2 view_lims(const sfk_view& v) {
    return make_float2(1.3f * v.img[0] / (2.0f * v.focal[0]), 1.3f * v.img[1] / (2.0f * v.focal[1]));
}

// Culls, EWA covariance, conic, mean2d, extent and tile bbox for one view,
// from the splat's view-independent parts (splat_rs, splat_cutoff) and the
// view's clamp limits (view_lims): the same operations as computing them
// inline, so a caller that projects one splat into many views computes those
// once.  Returns false when the splat is culled.
__device__ __forceinline__ bool project_geom_pre(const sfk_view& v, const SplatIn& s,
                                                 const float (&m)[9], float cutoff, float2 lim,
                                                 float inv_tile, ProjGeom& o) {
    if (s.opacity < kAlphaCutoff) return false;
    float cx = dot3f(v.rot0[0], v.rot0[1], v.rot0[2], s.mx, s.my, s.mz) + v.trans[0];
    float cy = dot3f(v.rot1[0], v.rot1[1], v.rot1[2], s.mx, s.my, s.mz) + v.trans[1];
    float cz = dot3f(v.rot2[0], v.rot2[1], v.rot2[2], s.mx, s.my, s.mz) + v.trans[2];
    if (cz <= 0.1f) return false;
    float iz = 1.0f / cz;
    // compute_cov2d (project.rs:175-221)
    const float m00 = m[0], m01 = m[1], m02 = m[2];
    const float m10 = m[3], m11 = m[4], m12 = m[5];
    const float m20 = m[6], m21 = m[7], m22 = m[8];
    const float lim_x = lim.x, lim_y = lim.y;
    float u = sfk_clampf(cx * iz, -lim_x, lim_x);
    float vv = sfk_clampf(cy * iz, -lim_y, lim_y);
    float fx_iz = v.focal[0] * iz;
    float fu_iz = fx_iz * u;
    float t00 = fx_iz * v.rot0[0] - fu_iz * v.rot2[0];
    float t01 = fx_iz * v.rot0[1] - fu_iz * v.rot2[1];
    float t02 = fx_iz * v.rot0[2] - fu_iz * v.rot2[2];
    float fy_iz = v.focal[1] * iz;
    float fv_iz = fy_iz * vv;
    float t10 = fy_iz * v.rot1[0] - fv_iz * v.rot2[0];
    float t11 = fy_iz * v.rot1[1] - fv_iz * v.rot2[1];
    float t12 = fy_iz * v.rot1[2] - fv_iz * v.rot2[2];
    float j00 = t00 * m00 + t01 * m10 + t02 * m20;
    float j01 = t00 * m01 + t01 * m11 + t02 * m21;
    float j02 = t00 * m02 + t01 * m12 + t02 * m22;
    float j10 = t10 * m00 + t11 * m10 + t12 * m20;
    float j11 = t10 * m01 + t11 * m11 + t12 * m21;
    float j12 = t10 * m02 + t11 * m12 + t12 * m22;
    float ca = dot3f(j00, j01, j02, j00, j01, j02) + 0.3f;
    float cb = dot3f(j00, j01, j02, j10, j11, j12);
    float cc = dot3f(j10, j11, j12, j10, j11, j12) + 0.3f;
    // compute_conic (project.rs:223-232)
    float det = fmaxf(ca * cc - cb * cb, 1e-6f);
    float inv_det = 1.0f / det;
    o.qa = cc * inv_det;
    o.qb = -cb * inv_det;
    o.qc = ca * inv_det;
    o.cz = cz;
    o.px = v.focal[0] * cx * iz + v.img[0] * 0.5f;
    o.py = v.focal[1] * cy * iz + v.img[1] * 0.5f;
    float ex = sqrtf(2.0f * cutoff * ca);
    float ey = sqrtf(2.0f * cutoff * cc);
    // tile_bbox (project.rs:37-46)
    float bx = v.tile_bounds[0], by = v.tile_bounds[1];
    uint32_t min_x = sfk_f2u(sfk_clampf((o.px - ex) * inv_tile, 0.0f, bx));
    uint32_t min_y = sfk_f2u(sfk_clampf((o.py - ey) * inv_tile, 0.0f, by));
    uint32_t max_x = sfk_f2u(sfk_clampf((o.px + ex) * inv_tile + 1.0f, 0.0f, bx));
    uint32_t max_y = sfk_f2u(sfk_clampf((o.py + ey) * inv_tile + 1.0f, 0.0f, by));
    o.count = (max_x - min_x) * (max_y - min_y);
    o.bb_lo = min_x | (min_y << 16);
    o.bb_hi = max_x | (max_y << 16);
    return true;
}

// project_geom_pre with every part computed here (one view).
__device__ __forceinline__ bool project_geom(const sfk_view& v, const SplatIn& s, float inv_tile,
                                             ProjGeom& o) {
    if (s.opacity < kAlphaCutoff) return false;
    float m[9];
    splat_rs(s, m);
    return project_geom_pre(v, s, m, splat_cutoff(s), view_lims(v), inv_tile, o);
}

// Tile bbox in one word, 8 bits per bound (min_x, min_y, max_x, max_y from
// the low byte up), for grids of at most 255 x 255 tiles; 0 for a culled
// splat (zero tiles).  It rides through the depth sort as the payload of
// the identity values, so the emission reads it in depth order.
// Derivatives D[k][j] of the SH basis functions of sh_to_rgb
// (project.rs:342-388, band 3 as written there) with respect to the unit
// view direction component j (x, y, z), coefficients k < CHS; the rest 0.
template <int CHS>
__device__ __forceinline__ void sh_basis_dir_grad(float x, float y, float z, float (&D)[16][3]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) D[k][0] = D[k][1] = D[k][2] = 0.f;
    constexpr float c1 = 0.48860252f, c20 = 0.9461747f, c22 = -1.0925485f, c23 = 0.54627424f;
    constexpr float k9 = -0.5900436f;
    if constexpr (CHS >= 4) {
        D[1][1] = -c1; D[2][2] = c1; D[3][0] = -c1;
    }
    if constexpr (CHS >= 9) {
        D[4][0] = 2.f * c23 * y; D[4][1] = 2.f * c23 * x;
        D[5][1] = c22 * z; D[5][2] = c22 * y;
        D[6][2] = 2.f * c20 * z;
        D[7][0] = c22 * z; D[7][2] = c22 * x;
        D[8][0] = 2.f * c23 * x; D[8][1] = -2.f * c23 * y;
    }
    if constexpr (CHS >= 16) {
        const float c1x = x * x - y * y, c1y = 2.f * x * y;
        const float tt0 = -2.285229f * z * z + 0.4570458f, tt1 = 1.4453057f * z;
        D[9][0] = k9 * c1y + k9 * x * 2.f * y + y * 2.f * x;
        D[9][1] = k9 * x * 2.f * x + c1x - 2.f * y * y;
        D[10][0] = k9 * c1x + k9 * x * 2.f * x - 2.f * y * y;
        D[10][1] = -k9 * x * 2.f * y - c1y - 2.f * x * y;
        D[11][1] = tt0; D[11][2] = -4.570458f * z * y;
        D[12][0] = tt0; D[12][2] = -4.570458f * z * x;
        D[13][2] = 5.5976451f * z * z - 1.119529f;
        D[14][0] = tt1 * 2.f * y; D[14][1] = tt1 * 2.f * x; D[14][2] = 1.4453057f * c1y;
        D[15][0] = tt1 * 2.f * x; D[15][1] = -tt1 * 2.f * y; D[15][2] = 1.4453057f * c1x;
    }
}

__device__ __forceinline__ uint32_t compact_bbox(const ProjGeom& o) {
    return (o.bb_lo & 0xffu) | ((o.bb_lo >> 16) << 8) | ((o.bb_hi & 0xffu) << 16) |
           ((o.bb_hi >> 16) << 24);
}

// Informational per-frame counts (isects, visible): reduced across the
// 256-thread workgroup, then added to one of kCountSlots slot pairs chosen
// by workgroup id.  One hot address for every workgroup serializes in L2
// (about 8K atomics per 1M splats); spread over 64 slot pairs it does not.
// slots[2*s] += isects, slots[2*s+1] += visible; readers sum the slots.
__device__ __forceinline__ void block_count_pair(uint32_t cnt, bool vis, uint32_t* slots) {
    __shared__ uint32_t s_c[2];
    if (threadIdx.x == 0) {
        s_c[0] = 0;
        s_c[1] = 0;
    }
    __syncthreads();
    uint32_t nvis = (uint32_t)__popcll(__ballot(vis));
    uint32_t ncnt = wave_sum_u32(cnt);
    if (lane_id() == 0) {
        if (ncnt) atomicAdd(&s_c[0], ncnt);
        if (nvis) atomicAdd(&s_c[1], nvis);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t* p = slots + 2 * (blockIdx.x % kCountSlots);
        if (s_c[0]) atomicAdd(&p[0], s_c[0]);
        if (s_c[1]) atomicAdd(&p[1], s_c[1]);
    }
}

}  // namespace sfk
