// Beta-distribution entropy in f64 on the device (src/seg/beta.rs:52-96):
// the reference evaluates it on the host to fill a 445x445 f32 table
// (beta.rs:100-144); here it is evaluated per splat in the EIG kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sfk {

// ln Gamma, Lanczos g = 7, 9 terms (beta.rs:55-70); x >= 1 here.
__device__ __forceinline__ double d_lgamma(double x) {
    const double z = x - 1.0;
    double series = 0.9999999999998099;
    series += 676.5203681218851 / (z + 1.0);
    series += -1259.1392167224028 / (z + 2.0);
    series += 771.3234287776531 / (z + 3.0);
    series += -176.6150291621406 / (z + 4.0);
    series += 12.507343278686905 / (z + 5.0);
    series += -0.13857109526572012 / (z + 6.0);
    series += 9.984369578019572e-6 / (z + 7.0);
    series += 1.5056327351493116e-7 / (z + 8.0);
    const double t = z + 7.5;
    return 0.91893853320467274178 + (z + 0.5) * log(t) - t + log(series);
}

// digamma: recurrence to x >= 6, then the asymptotic series (beta.rs:75-86).
__device__ __forceinline__ double d_digamma(double x) {
    double r = 0.0;
    while (x < 6.0) {
        r -= 1.0 / x;
        x += 1.0;
    }
    const double inv = 1.0 / x, inv2 = inv * inv;
    return r + log(x) - 0.5 * inv - inv2 / 12.0 + inv2 * inv2 / 120.0 -
           inv2 * inv2 * inv2 / 252.0;
}

// H(a, b) of Beta(a, b) (beta.rs:88-96).
__device__ __forceinline__ double d_beta_entropy(double a, double b) {
    const double s = a + b;
    return d_lgamma(a) + d_lgamma(b) - d_lgamma(s) - (a - 1.0) * d_digamma(a) -
           (b - 1.0) * d_digamma(b) + (s - 2.0) * d_digamma(s);
}

// max(0, H0 - H(a k, b k)), k = 1 + eps/(a+b), eps = f32(q)/scale; 0 when
// eps = 0 (beta.rs:365-415).  H0 = H(a, b).
__device__ __forceinline__ double d_entropy_drop(float a, float b, double h0, uint32_t q,
                                                 float scale) {
    float eps = (float)q / scale;
    if (!(eps > 0.0f)) return 0.0;
    double ai = a, bi = b;
    double k = 1.0 + (double)eps / (ai + bi);
    double d = h0 - d_beta_entropy(ai * k, bi * k);
    return d > 0.0 ? d : 0.0;
}

}  // namespace sfk
