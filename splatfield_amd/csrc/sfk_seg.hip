// B3-Seg kernels for gfx950 (src/seg/mod.rs, src/seg/beta.rs): buffer fill,
// analytic EIG with the Beta entropy in f64 on the device (in place of the
// 445x445 f32 table, beta.rs:100-144), and the conjugate Beta update.
#include "sfk_device.h"
#include "sfk_entropy.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ buf, uint32_t v,
                                                  size_t count) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < count; i += stride) buf[i] = v;
}

// Two buffers zeroed in one launch (a frame's counters and look-back words).
__global__ __launch_bounds__(256) void k_zero2_u32(uint32_t* __restrict__ a, size_t na,
                                                   uint32_t* __restrict__ b, size_t nb) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < na + nb; i += stride) {
        if (i < na) a[i] = 0u;
        else b[i - na] = 0u;
    }
}

// Plain streaming copy (a known-traffic kernel for counter calibration).
__global__ __launch_bounds__(256) void k_copy_u32(uint4* __restrict__ dst,
                                                  const uint4* __restrict__ src, size_t count4) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < count4; i += stride) dst[i] = src[i];
}

constexpr uint32_t kEigThreads = 256;
constexpr uint32_t kEigEpt = 4;
constexpr uint32_t kEigBlock = kEigThreads * kEigEpt;

// Per-splat dH = max(0, H(a,b) - H(a k, b k)), k = 1 + eps/(a+b),
// eps = f32(q)/scale, skipped at eps = 0 (beta.rs:172-186); per-block f64 sums.
__global__ __launch_bounds__(kEigThreads) void k_eig_partial(const float* __restrict__ a,
                                                             const float* __restrict__ b,
                                                             const uint32_t* __restrict__ resp,
                                                             uint32_t n, float scale,
                                                             double* __restrict__ partials) {
    __shared__ double s_w[kEigThreads / 64];
    double sum = 0.0;
    const size_t base = (size_t)blockIdx.x * kEigBlock + threadIdx.x;
#pragma unroll
    for (uint32_t e = 0; e < kEigEpt; ++e) {
        size_t i = base + (size_t)e * kEigThreads;
        if (i < n) {
            float eps = (float)resp[i] / scale;
            if (eps > 0.0f) {
                double ai = a[i], bi = b[i];
                double k = 1.0 + (double)eps / (ai + bi);
                double d = d_beta_entropy(ai, bi) - d_beta_entropy(ai * k, bi * k);
                sum += d > 0.0 ? d : 0.0;
            }
        }
    }
    sum = wave_sum_f64(sum);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t w = 0; w < kEigThreads / 64; ++w) t += s_w[w];
        partials[blockIdx.x] = t;
    }
}

// Deterministic final sum of the block partials (fixed order).
__global__ __launch_bounds__(256) void k_eig_final(const double* __restrict__ partials,
                                                   uint32_t np, double* __restrict__ out) {
    __shared__ double s_w[4];
    double sum = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += 256) sum += partials[i];
    sum = wave_sum_f64(sum);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) *out = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// beta_update_kernel (beta.rs:203-212).
__global__ __launch_bounds__(256) void k_beta_update(float* __restrict__ a, float* __restrict__ b,
                                                     const uint32_t* __restrict__ fg,
                                                     const uint32_t* __restrict__ bg, uint32_t n,
                                                     float scale) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) {
        a[i] += (float)fg[i] / scale;
        b[i] += (float)bg[i] / scale;
    }
}

// tint_colors (src/seg/paint.rs:113-122): DC channels of the splats with
// a > b (ties stay) set to (r, g, bl); DC channel k of splat i at k*n + i.
__global__ __launch_bounds__(256) void k_tint_colors(float* __restrict__ sh,
                                                     const float* __restrict__ a,
                                                     const float* __restrict__ b, uint32_t n,
                                                     float r, float g, float bl) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n && a[i] > b[i]) {
        sh[i] = r;
        sh[(size_t)n + i] = g;
        sh[2 * (size_t)n + i] = bl;
    }
}

// copy_f32 (src/seg/paint.rs:128-134) over a known count.
__global__ __launch_bounds__(256) void k_copy_f32(float* __restrict__ dst,
                                                  const float* __restrict__ src, size_t count) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < count; i += stride) dst[i] = src[i];
}

// views::localize (views.rs:32-67) on the device: posterior-weighted centre
// and mean distance of the splats the MAP labels foreground (a > b, weight
// m = a / (a + b)).  Pass 0 sums (m, m x, m y, m z); pass 1 sums m |p - c|
// with c from pass 0.  f64 partials per block, summed in block order by the
// last launch of each pass (deterministic; the reference's sequential f32
// loop carries its own rounding).
constexpr uint32_t kLocThreads = 256;

template <int PASS>
__global__ __launch_bounds__(kLocThreads) void k_localize(const float* __restrict__ A,
                                                          const float* __restrict__ a,
                                                          const float* __restrict__ b, uint32_t n,
                                                          const double* __restrict__ center,
                                                          double* __restrict__ partials) {
    constexpr int NV = PASS == 0 ? 4 : 1;
    __shared__ double s_w[NV][kLocThreads / 64];
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    const size_t N = n;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if constexpr (PASS == 1) {
        cx = center[0]; cy = center[1]; cz = center[2];
    }
    for (uint32_t i = blockIdx.x * kLocThreads + threadIdx.x; i < n;
         i += gridDim.x * kLocThreads) {
        const float ai = a[i], bi = b[i];
        if (!(ai > bi)) continue;
        const double m = (double)(ai / (ai + bi));
        const double x = A[i], y = A[N + i], z = A[2 * N + i];
        if constexpr (PASS == 0) {
            acc[0] += m;
            acc[1] += m * x;
            acc[2] += m * y;
            acc[3] += m * z;
        } else {
            const double dx = x - cx, dy = y - cy, dz = z - cz;
            acc[0] += m * sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double v = wave_sum_f64(acc[k]);
        if (lane_id() == 0) s_w[k][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)NV) {
        double t = 0.0;
        for (uint32_t w = 0; w < kLocThreads / 64; ++w) t += s_w[threadIdx.x][w];
        partials[(size_t)blockIdx.x * NV + threadIdx.x] = t;
    }
}

// Fixed-order reduction of the partials; pass 0 turns (sw, sx, sy, sz) into
// the centre, pass 1 writes the radius.  out = [cx, cy, cz, sw, radius].
template <int PASS>
__global__ void k_localize_final(const double* __restrict__ partials, uint32_t nb,
                                 double* __restrict__ out) {
    constexpr int NV = PASS == 0 ? 4 : 1;
    if (threadIdx.x != 0) return;
    double t[NV];
    for (int k = 0; k < NV; ++k) t[k] = 0.0;
    for (uint32_t i = 0; i < nb; ++i)
        for (int k = 0; k < NV; ++k) t[k] += partials[(size_t)i * NV + k];
    if constexpr (PASS == 0) {
        const double sw = t[0];
        out[3] = sw;
        out[0] = sw > 0.0 ? t[1] / sw : 0.0;
        out[1] = sw > 0.0 ? t[2] / sw : 0.0;
        out[2] = sw > 0.0 ? t[3] / sw : 0.0;
    } else {
        out[4] = out[3] > 0.0 ? t[0] / out[3] : 0.0;
    }
}

}  // namespace

uint32_t localize_blocks(uint32_t n) {
    const uint32_t nb = div_up(n ? n : 1, kLocThreads * 16u);
    return nb < 1024u ? nb : 1024u;
}

hipError_t launch_localize(hipStream_t s, const float* attrs, const float* a, const float* b,
                           uint32_t n, double* partials, double* out) {
    const uint32_t nb = localize_blocks(n);
    k_localize<0><<<nb, kLocThreads, 0, s>>>(attrs, a, b, n, nullptr, partials);
    k_localize_final<0><<<1, 64, 0, s>>>(partials, nb, out);
    k_localize<1><<<nb, kLocThreads, 0, s>>>(attrs, a, b, n, out, partials);
    k_localize_final<1><<<1, 64, 0, s>>>(partials, nb, out);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t* buf, uint32_t v, size_t count) {
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_fill_u32<<<(uint32_t)blocks, 256, 0, s>>>(buf, v, count);
    return hipGetLastError();
}

hipError_t launch_zero2_u32(hipStream_t s, uint32_t* a, size_t na, uint32_t* b, size_t nb) {
    if (na + nb == 0) return hipSuccess;
    size_t blocks = (na + nb + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_zero2_u32<<<(uint32_t)blocks, 256, 0, s>>>(a, na, b, nb);
    return hipGetLastError();
}

hipError_t launch_copy_u32(hipStream_t s, uint32_t* dst, const uint32_t* src, size_t count) {
    size_t c4 = count / 4;
    if (c4 == 0) return hipSuccess;
    size_t blocks = (c4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_copy_u32<<<(uint32_t)blocks, 256, 0, s>>>(reinterpret_cast<uint4*>(dst),
                                                reinterpret_cast<const uint4*>(src), c4);
    return hipGetLastError();
}

uint32_t eig_num_partials(uint32_t n) { return div_up(n > 0 ? n : 1, kEigBlock); }

hipError_t launch_eig(hipStream_t s, const float* a, const float* b, const uint32_t* resp,
                      uint32_t n, float scale, double* partials, double* out) {
    uint32_t np = eig_num_partials(n);
    if (n > 0) k_eig_partial<<<np, kEigThreads, 0, s>>>(a, b, resp, n, scale, partials);
    else np = 0;
    k_eig_final<<<1, 256, 0, s>>>(partials, np, out);
    return hipGetLastError();
}

hipError_t launch_beta_update(hipStream_t s, float* a, float* b, const uint32_t* fg,
                              const uint32_t* bg, uint32_t n, float scale) {
    if (n == 0) return hipSuccess;
    k_beta_update<<<div_up(n, 256), 256, 0, s>>>(a, b, fg, bg, n, scale);
    return hipGetLastError();
}

hipError_t launch_tint_colors(hipStream_t s, float* sh, const float* a, const float* b,
                              uint32_t n, float r, float g, float bl) {
    if (n == 0) return hipSuccess;
    k_tint_colors<<<div_up(n, 256), 256, 0, s>>>(sh, a, b, n, r, g, bl);
    return hipGetLastError();
}

hipError_t launch_copy_f32(hipStream_t s, float* dst, const float* src, size_t count) {
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_copy_f32<<<(uint32_t)blocks, 256, 0, s>>>(dst, src, count);
    return hipGetLastError();
}

}  // namespace sfk
