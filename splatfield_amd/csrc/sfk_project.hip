// project_splats for gfx950 (src/project.rs:234-388, docs/01_project.md).
//
// One thread per splat over coalesced field-major planes.  Differences from
// the reference kernel, all deliberate:
//  * no compaction atomic: every splat i writes its depth key at index i
//    (0xffffffff when culled), so the stable depth sort that follows ranks
//    visible splats by (depth, index) — a deterministic tie order;
//  * the two global counters are reduced per wavefront first (one atomic per
//    wave instead of two per splat);
//  * the projected row is 12 floats (three 16-byte stores/loads):
//    [mx, my, conic_a, conic_b, conic_c, opacity, r, g, b, 0, 0, 0].
#include "sfk_project_core.h"

namespace sfk {

namespace {

constexpr float kShC0 = 0.2820948f;  // project.rs:329
constexpr float kShC1 = 0.48860252f;
constexpr float kShC20 = 0.9461747f;
constexpr float kShC21 = 0.31539157f;
constexpr float kShC22 = -1.0925485f;
constexpr float kShC23 = 0.54627424f;

// The splat's SH coefficients, plane 3k + c (coefficient k, channel c),
// all loads issued before any use.
template <int CHS>
__device__ __forceinline__ void sh_load(const float* __restrict__ shs, uint32_t s, uint32_t n,
                                        float (&c)[3 * CHS]) {
#pragma unroll
    for (int k = 0; k < 3 * CHS; ++k) c[k] = __builtin_nontemporal_load(shs + (size_t)k * n + s);
}

// sh_to_rgb (project.rs:342-388) with its band-3 basis as written.
template <int CHS>
__device__ __forceinline__ void sh_to_rgb(float x, float y, float z, const float (&SH)[3 * CHS],
                                          float& r, float& g, float& b) {
    r = kShC0 * SH[0];
    g = kShC0 * SH[1];
    b = kShC0 * SH[2];
    if constexpr (CHS >= 4) {
        r += kShC1 * ((-y * SH[3] + z * SH[6]) - x * SH[9]);
        g += kShC1 * ((-y * SH[4] + z * SH[7]) - x * SH[10]);
        b += kShC1 * ((-y * SH[5] + z * SH[8]) - x * SH[11]);
    }
    if constexpr (CHS >= 9) {
        float b4 = kShC22 * z * x;
        float b5 = kShC22 * z * y;
        float b6 = kShC20 * z * z - kShC21;
        float b7 = kShC23 * 2.0f * x * y;
        float b8 = kShC23 * (x * x - y * y);
        r += b6 * SH[18] + b7 * SH[12] + b5 * SH[15] + b4 * SH[21] + b8 * SH[24];
        g += b6 * SH[19] + b7 * SH[13] + b5 * SH[16] + b4 * SH[22] + b8 * SH[25];
        b += b6 * SH[20] + b7 * SH[14] + b5 * SH[17] + b4 * SH[23] + b8 * SH[26];
    }
    if constexpr (CHS >= 16) {
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
        r += b9 * SH[27] + b10 * SH[30] + b11 * SH[33] + b12 * SH[36] + b13 * SH[39] +
             b14 * SH[42] + b15 * SH[45];
        g += b9 * SH[28] + b10 * SH[31] + b11 * SH[34] + b12 * SH[37] + b13 * SH[40] +
             b14 * SH[43] + b15 * SH[46];
        b += b9 * SH[29] + b10 * SH[32] + b11 * SH[35] + b12 * SH[38] + b13 * SH[41] +
             b14 * SH[44] + b15 * SH[47];
    }
}

// CHS = SH coefficients per channel (1, 4, 9, 16); CHS = 0 skips colour
// (geometry-only projection for the responsibility modes); CHS = -1 paints
// the posterior mean m = a/(a+b) as grayscale with no +0.5 offset
// (paint_posterior, src/seg/paint.rs:99-109, fused here: no SH is read).
template <int CHS, bool JAC>
__global__ __launch_bounds__(256) void k_project(sfk_view v, const float* __restrict__ A,
                                                 const float* __restrict__ shs, uint32_t n,
                                                 float inv_tile,
                                                 uint32_t* __restrict__ depth_keys,
                                                 float4* __restrict__ proj,
                                                 uint32_t* __restrict__ tile_counts,
                                                 uint2* __restrict__ bbox,
                                                 uint32_t* __restrict__ cbbox,
                                                 uint32_t* __restrict__ counters,
                                                 float* __restrict__ dcol,
                                                 const float* __restrict__ post_a,
                                                 const float* __restrict__ post_b) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool vis = false;
    uint32_t cnt = 0, cb = 0;
    if (i < n) {
        SplatIn s = load_splat(A, i, n);
        ProjGeom o;
        uint32_t key = 0xffffffffu;
        if (project_geom(v, s, inv_tile, o)) {
            vis = true;
            key = __builtin_bit_cast(uint32_t, o.cz);
            cnt = o.count;
            float cr = 0.0f, cg = 0.0f, cbl = 0.0f;
            if constexpr (CHS < 0) {
                const float pa = post_a[i];
                const float m = pa / (pa + post_b[i]);
                cr = cg = cbl = m;
            }
            if constexpr (CHS > 0) {
                float dx = s.mx - v.world_pos[0], dy = s.my - v.world_pos[1],
                      dz = s.mz - v.world_pos[2];
                float inv = 1.0f / sqrtf(dot3f(dx, dy, dz, dx, dy, dz));
                float shc[3 * CHS];
                sh_load<CHS>(shs, i, n, shc);
                sh_to_rgb<CHS>(dx * inv, dy * inv, dz * inv, shc, cr, cg, cbl);
                if constexpr (JAC) {
                    // d colour / d unit direction for the adjoint:
                    // J[c][j] = sum_k sh[3k + c] D[k][j] (plane 3c + j).
                    float D[16][3];
                    sh_basis_dir_grad<CHS>(dx * inv, dy * inv, dz * inv, D);
                    float J[3][3] = {};
#pragma unroll
                    for (int k = 1; k < CHS; ++k)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int j = 0; j < 3; ++j) J[c][j] += shc[3 * k + c] * D[k][j];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int j = 0; j < 3; ++j) dcol[(size_t)(3 * c + j) * n + i] = J[c][j];
                }
            }
            float4* row = proj + (size_t)i * 3;
            row[0] = make_float4(o.px, o.py, o.qa, o.qb);
            constexpr float kOff = CHS < 0 ? 0.0f : 0.5f;  // project.rs:306-310
            row[1] = make_float4(o.qc, s.opacity, cr + kOff, cg + kOff);
            row[2] = make_float4(cbl + kOff, 0.0f, 0.0f, 0.0f);
            bbox[i] = make_uint2(o.bb_lo, o.bb_hi);
            cb = compact_bbox(o);
        }
        depth_keys[i] = key;
        tile_counts[i] = cnt;
        if (cbbox) cbbox[i] = cb;
    }
    block_count_pair(cnt, vis, counters);
}

}  // namespace

hipError_t launch_project(hipStream_t s, const sfk_view& v, const float* attrs, const float* sh,
                          uint32_t n, uint32_t sh_per_ch, uint32_t tile_w, uint32_t* depth_keys,
                          float* proj, uint32_t* tile_counts, uint32_t* bbox, uint32_t* counters,
                          uint32_t* cbbox, float* dcol, const float* post_a,
                          const float* post_b) {
    if (n == 0) return hipSuccess;
    dim3 grid(div_up(n, 256)), block(256);
    float inv_tile = 1.0f / (float)tile_w;
    auto* p4 = reinterpret_cast<float4*>(proj);
    auto* bb = reinterpret_cast<uint2*>(bbox);
#define SFK_PROJ(C)                                                                        \
    do {                                                                                   \
        if (dcol)                                                                          \
            k_project<C, true><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4, \
                                                      tile_counts, bb, cbbox, counters, dcol,  \
                                                      nullptr, nullptr);                       \
        else                                                                               \
            k_project<C, false><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys,   \
                                                       p4, tile_counts, bb, cbbox, counters,   \
                                                       nullptr, nullptr, nullptr);             \
    } while (0)
    if (post_a || post_b) {
        if (!post_a || !post_b || dcol) return hipErrorInvalidValue;
        k_project<-1, false><<<grid, block, 0, s>>>(v, attrs, nullptr, n, inv_tile, depth_keys, p4,
                                                    tile_counts, bb, cbbox, counters, nullptr,
                                                    post_a, post_b);
        return hipGetLastError();
    }
    switch (sh_per_ch) {
        case 0: SFK_PROJ(0); break;
        case 1: SFK_PROJ(1); break;
        case 4: SFK_PROJ(4); break;
        case 9: SFK_PROJ(9); break;
        case 16: SFK_PROJ(16); break;
        default: return hipErrorInvalidValue;
    }
#undef SFK_PROJ
    return hipGetLastError();
}

}  // namespace sfk
