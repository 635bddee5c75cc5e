// project_splats for gfx950 (src/project.rs:295-388, docs/01_project.md).
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
#include "sfk_device.h"
#include "sfk_internal.h"
#include "sfk_math.h"

namespace sfk {

namespace {

constexpr float kAlphaCutoff = 10.0f / 255.0f;  // project.rs:74
constexpr float kShC0 = 0.2820948f;              // project.rs:390-401
constexpr float kShC1 = 0.48860252f;
constexpr float kShC20 = 0.9461747f;
constexpr float kShC21 = 0.31539157f;
constexpr float kShC22 = -1.0925485f;
constexpr float kShC23 = 0.54627424f;

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// sh_to_rgb (project.rs:403-449) with its band-3 basis as written.
template <int CHS>
__device__ __forceinline__ void sh_to_rgb(float x, float y, float z, uint32_t s, uint32_t n,
                                          const float* __restrict__ shs, float& r, float& g,
                                          float& b) {
    const size_t st = n;
    auto SH = [&](int k) { return __builtin_nontemporal_load(shs + (size_t)k * st + s); };
    r = kShC0 * SH(0);
    g = kShC0 * SH(1);
    b = kShC0 * SH(2);
    if constexpr (CHS >= 4) {
        r += kShC1 * ((-y * SH(3) + z * SH(6)) - x * SH(9));
        g += kShC1 * ((-y * SH(4) + z * SH(7)) - x * SH(10));
        b += kShC1 * ((-y * SH(5) + z * SH(8)) - x * SH(11));
    }
    if constexpr (CHS >= 9) {
        float b4 = kShC22 * z * x;
        float b5 = kShC22 * z * y;
        float b6 = kShC20 * z * z - kShC21;
        float b7 = kShC23 * 2.0f * x * y;
        float b8 = kShC23 * (x * x - y * y);
        r += b6 * SH(18) + b7 * SH(12) + b5 * SH(15) + b4 * SH(21) + b8 * SH(24);
        g += b6 * SH(19) + b7 * SH(13) + b5 * SH(16) + b4 * SH(22) + b8 * SH(25);
        b += b6 * SH(20) + b7 * SH(14) + b5 * SH(17) + b4 * SH(23) + b8 * SH(26);
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
        r += b9 * SH(27) + b10 * SH(30) + b11 * SH(33) + b12 * SH(36) + b13 * SH(39) +
             b14 * SH(42) + b15 * SH(45);
        g += b9 * SH(28) + b10 * SH(31) + b11 * SH(34) + b12 * SH(37) + b13 * SH(40) +
             b14 * SH(43) + b15 * SH(46);
        b += b9 * SH(29) + b10 * SH(32) + b11 * SH(35) + b12 * SH(38) + b13 * SH(41) +
             b14 * SH(44) + b15 * SH(47);
    }
}

// CHS = SH coefficients per channel (1, 4, 9, 16); CHS = 0 skips colour
// (geometry-only projection for the responsibility modes).
template <int CHS>
__global__ __launch_bounds__(256) void k_project(sfk_view v, const float* __restrict__ A,
                                                 const float* __restrict__ shs, uint32_t n,
                                                 float inv_tile,
                                                 uint32_t* __restrict__ depth_keys,
                                                 float4* __restrict__ proj,
                                                 uint32_t* __restrict__ tile_counts,
                                                 uint2* __restrict__ bbox,
                                                 uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool vis = false;
    uint32_t cnt = 0;
    if (i < n) {
        const size_t N = n;
        float mx = A[0 * N + i], my = A[1 * N + i], mz = A[2 * N + i];
        float qw = A[3 * N + i], qx = A[4 * N + i], qy = A[5 * N + i], qz = A[6 * N + i];
        float sx = sfk_expf(A[7 * N + i]), sy = sfk_expf(A[8 * N + i]),
              sz = sfk_expf(A[9 * N + i]);
        float opacity = 1.0f / (1.0f + sfk_expf(-A[10 * N + i]));
        uint32_t key = 0xffffffffu;
        if (!(opacity < kAlphaCutoff)) {
            float cx = dot3(v.rot0[0], v.rot0[1], v.rot0[2], mx, my, mz) + v.trans[0];
            float cy = dot3(v.rot1[0], v.rot1[1], v.rot1[2], mx, my, mz) + v.trans[1];
            float cz = dot3(v.rot2[0], v.rot2[1], v.rot2[2], mx, my, mz) + v.trans[2];
            if (!(cz <= 0.1f)) {
                float iz = 1.0f / cz;
                // compute_cov2d (project.rs:236-282)
                float x2 = qx * qx, y2 = qy * qy, z2 = qz * qz;
                float xy = qx * qy, xz = qx * qz, yz = qy * qz;
                float wx = qw * qx, wy = qw * qy, wz = qw * qz;
                float r00 = 1.0f - 2.0f * (y2 + z2), r01 = 2.0f * (xy - wz), r02 = 2.0f * (xz + wy);
                float r10 = 2.0f * (xy + wz), r11 = 1.0f - 2.0f * (x2 + z2), r12 = 2.0f * (yz - wx);
                float r20 = 2.0f * (xz - wy), r21 = 2.0f * (yz + wx), r22 = 1.0f - 2.0f * (x2 + y2);
                float m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
                float m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
                float m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
                float lim_x = 1.3f * v.img[0] / (2.0f * v.focal[0]);
                float lim_y = 1.3f * v.img[1] / (2.0f * v.focal[1]);
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
                float ca = dot3(j00, j01, j02, j00, j01, j02) + 0.3f;
                float cb = dot3(j00, j01, j02, j10, j11, j12);
                float cc = dot3(j10, j11, j12, j10, j11, j12) + 0.3f;
                // compute_conic (project.rs:284-293)
                float det = fmaxf(ca * cc - cb * cb, 1e-6f);
                float inv_det = 1.0f / det;

                vis = true;
                key = __builtin_bit_cast(uint32_t, cz);

                float cr = 0.0f, cg = 0.0f, cbl = 0.0f;
                if constexpr (CHS > 0) {
                    float dx = mx - v.world_pos[0], dy = my - v.world_pos[1],
                          dz = mz - v.world_pos[2];
                    float inv = 1.0f / sqrtf(dot3(dx, dy, dz, dx, dy, dz));
                    sh_to_rgb<CHS>(dx * inv, dy * inv, dz * inv, i, n, shs, cr, cg, cbl);
                }
                float px = v.focal[0] * cx * iz + v.img[0] * 0.5f;
                float py = v.focal[1] * cy * iz + v.img[1] * 0.5f;
                float4* row = proj + (size_t)i * 3;
                row[0] = make_float4(px, py, cc * inv_det, -cb * inv_det);
                row[1] = make_float4(ca * inv_det, opacity, cr + 0.5f, cg + 0.5f);
                row[2] = make_float4(cbl + 0.5f, 0.0f, 0.0f, 0.0f);

                float cutoff = sfk_logf(opacity / kAlphaCutoff);
                float ex = sqrtf(2.0f * cutoff * ca);
                float ey = sqrtf(2.0f * cutoff * cc);
                // tile_bbox (project.rs:98-107)
                float bx = v.tile_bounds[0], by = v.tile_bounds[1];
                uint32_t min_x = sfk_f2u(sfk_clampf((px - ex) * inv_tile, 0.0f, bx));
                uint32_t min_y = sfk_f2u(sfk_clampf((py - ey) * inv_tile, 0.0f, by));
                uint32_t max_x = sfk_f2u(sfk_clampf((px + ex) * inv_tile + 1.0f, 0.0f, bx));
                uint32_t max_y = sfk_f2u(sfk_clampf((py + ey) * inv_tile + 1.0f, 0.0f, by));
                cnt = (max_x - min_x) * (max_y - min_y);
                bbox[i] = make_uint2(min_x | (min_y << 16), max_x | (max_y << 16));
            }
        }
        depth_keys[i] = key;
        tile_counts[i] = cnt;
    }
    uint32_t nvis = (uint32_t)__popcll(__ballot(vis));
    uint32_t ncnt = wave_sum_u32(cnt);
    if (lane_id() == 0) {
        if (ncnt) atomicAdd(&counters[0], ncnt);
        if (nvis) atomicAdd(&counters[1], nvis);
    }
}

}  // namespace

hipError_t launch_project(hipStream_t s, const sfk_view& v, const float* attrs, const float* sh,
                          uint32_t n, uint32_t sh_per_ch, uint32_t tile_w, uint32_t* depth_keys,
                          float* proj, uint32_t* tile_counts, uint32_t* bbox, uint32_t* counters) {
    if (n == 0) return hipSuccess;
    dim3 grid(div_up(n, 256)), block(256);
    float inv_tile = 1.0f / (float)tile_w;
    auto* p4 = reinterpret_cast<float4*>(proj);
    auto* bb = reinterpret_cast<uint2*>(bbox);
    switch (sh_per_ch) {
        case 0:
            k_project<0><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4,
                                                 tile_counts, bb, counters);
            break;
        case 1:
            k_project<1><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4,
                                                 tile_counts, bb, counters);
            break;
        case 4:
            k_project<4><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4,
                                                 tile_counts, bb, counters);
            break;
        case 9:
            k_project<9><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4,
                                                 tile_counts, bb, counters);
            break;
        case 16:
            k_project<16><<<grid, block, 0, s>>>(v, attrs, sh, n, inv_tile, depth_keys, p4,
                                                  tile_counts, bb, counters);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sfk
