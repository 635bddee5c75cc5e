// Scene ingestion from SOG archives (src/sog.rs:97-198).  The host unzips
// and decodes the WebP sheets into RGBA8 (as the reference's image crate
// does); one kernel then decodes every splat from its texels: 16-bit
// log-space means, codebook scales / DC / opacity, smallest-three
// quaternions and the palette-coded higher SH bands.
#include <math.h>
#include <string.h>

#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

struct SogArgs {
    const uint32_t *lo, *hi, *scales, *quats, *sh0, *labels, *centroids;
    const float* cb;  // 768 floats: scales, sh0, shN codebooks
    float* attrs;
    float* sh;
    uint32_t* status;  // [first bad rotation tag splat, first bad label splat]
    uint64_t n, centroid_px;
    float mins[3], spans[3];
    uint32_t sh_per_ch, centroid_w, palette_count;
};

// sog.rs:66-68: sign(v) * expm1(|v|), the exponential in f64 then rounded.
__device__ inline float inv_log(float v) {
    return copysignf(1.0f, v) * (float)expm1((double)fabsf(v));
}

// sog.rs:70-73: ln(e / (1 - e)), e clamped to [1e-6, 1 - 1e-6].
__device__ inline float logit(float y) {
    const float e = fminf(fmaxf(y, 1e-6f), 1.0f - 1e-6f);
    return (float)log((double)(e / (1.0f - e)));
}

// sog.rs:80-91: three components in [-1/sqrt2, 1/sqrt2], the fourth (at
// tag - 252 in w, x, y, z order) recovered from the unit norm.
__device__ inline float unpack_u(uint32_t v) {
    return ((float)v / 255.0f * 2.0f - 1.0f) / 1.41421356f;
}

__global__ __launch_bounds__(256) void k_sog_decode(SogArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t n = a.n;
    const uint32_t lo = a.lo[i], hi = a.hi[i], sc = a.scales[i], qt = a.quats[i],
                   c0 = a.sh0[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t q16 = ((lo >> (8 * k)) & 0xffu) | (((hi >> (8 * k)) & 0xffu) << 8);
        const float t = (float)q16 / 65535.0f;
        a.attrs[(uint64_t)k * n + i] = inv_log(a.mins[k] + a.spans[k] * t);
        a.attrs[(uint64_t)(7 + k) * n + i] = a.cb[(sc >> (8 * k)) & 0xffu];
        a.sh[(uint64_t)k * n + i] = a.cb[256 + ((c0 >> (8 * k)) & 0xffu)];
    }
    a.attrs[10 * n + i] = logit((float)(c0 >> 24) / 255.0f);
    const uint32_t tag = qt >> 24;
    float q[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (tag >= 252) {
        const float x = unpack_u(qt & 0xffu), y = unpack_u((qt >> 8) & 0xffu),
                    z = unpack_u((qt >> 16) & 0xffu);
        const float d = sqrtf(fmaxf(1.0f - x * x - y * y - z * z, 0.0f));
        const uint32_t slot = tag - 252;  // 252: w omitted ... 255: z omitted
        q[0] = slot == 0 ? d : x;
        q[1] = slot == 0 ? x : slot == 1 ? d : y;
        q[2] = slot <= 1 ? y : slot == 2 ? d : z;
        q[3] = slot <= 2 ? z : d;
    } else {
        atomicMin(&a.status[0], (uint32_t)i);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) a.attrs[(uint64_t)(3 + k) * n + i] = q[k];
    if (a.sh_per_ch > 1) {
        const uint32_t cpc = a.sh_per_ch - 1;
        const uint32_t lb = a.labels[i];
        const uint32_t label = (lb & 0xffu) | (lb & 0xff00u);
        const uint32_t per_row = a.centroid_w / cpc;
        // sog.rs:201-205, the palette's first texel in the sheet.
        const uint64_t first = per_row ? (uint64_t)(label / per_row) * a.centroid_w +
                                             (uint64_t)(label % per_row) * cpc
                                       : 0;
        const bool ok = label < a.palette_count && per_row && first + cpc <= a.centroid_px;
        if (!ok) atomicMin(&a.status[1], (uint32_t)i);
        for (uint32_t j = 0; j < cpc; ++j) {
            const uint32_t px = ok ? a.centroids[first + j] : 0u;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                a.sh[(uint64_t)((j + 1) * 3 + k) * n + i] =
                    ok ? a.cb[512 + ((px >> (8 * k)) & 0xffu)] : 0.0f;
        }
    }
}

}  // namespace

}  // namespace sfk

using namespace sfk;

extern "C" int sfk_sog_decode(sfk_ctx* c, const sfk_sog_meta* m, const uint8_t* means_lo,
                              const uint8_t* means_hi, const uint8_t* scales,
                              const uint8_t* quats, const uint8_t* sh0, const uint8_t* labels,
                              const uint8_t* centroids, const float* codebooks, float* attrs,
                              float* sh, uint32_t* status) {
    if (!c || !m || !means_lo || !means_hi || !scales || !quats || !sh0 || !codebooks ||
        !attrs || !sh || !status)
        return SFK_E_ARG;
    if (m->n == 0 || m->n > 0xffffffffull)
        return fail(c, SFK_E_ARG, "SOG with %llu splats", (unsigned long long)m->n);
    if (!(m->sh_per_ch == 1 || m->sh_per_ch == 4 || m->sh_per_ch == 9 || m->sh_per_ch == 16))
        return fail(c, SFK_E_ARG, "SOG: %u SH coefficients per channel", m->sh_per_ch);
    if (m->sh_per_ch > 1 && (!labels || !centroids || m->centroid_w == 0))
        return fail(c, SFK_E_ARG, "SOG: shN sheets missing");
    SogArgs a;
    memset(&a, 0, sizeof a);
    a.lo = reinterpret_cast<const uint32_t*>(means_lo);
    a.hi = reinterpret_cast<const uint32_t*>(means_hi);
    a.scales = reinterpret_cast<const uint32_t*>(scales);
    a.quats = reinterpret_cast<const uint32_t*>(quats);
    a.sh0 = reinterpret_cast<const uint32_t*>(sh0);
    a.labels = reinterpret_cast<const uint32_t*>(labels);
    a.centroids = reinterpret_cast<const uint32_t*>(centroids);
    a.cb = codebooks;
    a.attrs = attrs;
    a.sh = sh;
    a.status = status;
    a.n = m->n;
    a.centroid_px = m->centroid_px;
    for (int k = 0; k < 3; ++k) {
        a.mins[k] = m->mins[k];
        a.spans[k] = m->maxs[k] - m->mins[k];  // glam Vec3 subtraction
    }
    a.sh_per_ch = m->sh_per_ch;
    a.centroid_w = m->centroid_w;
    a.palette_count = m->palette_count;
    SFK_PROF(c, "sog_decode");
    SFK_HIP(c, launch_fill_u32(c->stream, status, 0xffffffffu, 2));
    k_sog_decode<<<(uint32_t)((m->n + 255) / 256), 256, 0, c->stream>>>(a);
    SFK_HIP(c, hipGetLastError());
    return SFK_OK;
}
