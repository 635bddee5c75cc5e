// Beta-distribution entropy in f64 on the device (src/seg/beta.rs:52-96):
// the reference evaluates it on the host to fill a 445x445 f32 table
// (beta.rs:100-144); here it is evaluated per splat in the EIG kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sfk {

// ln Gamma(x) and digamma(x) together, for x >= 1, in f64.  The reference
// evaluates them separately (Lanczos g = 7 and a recurrence to x >= 6 with a
// short asymptotic series, beta.rs:55-86) to fill its f32 table; here both
// share one shift to z = x + 8 >= 9:
//   ln Gamma(x) = ln Gamma(z) - ln P,   psi(x) = psi(z) - P'/P,
//   P = prod_{i<8} (x + i)  (P and P' by one product recurrence),
// then Stirling's series for ln Gamma(z) and the asymptotic series for
// psi(z), each to the 1/z^9 (resp. 1/z^10) Bernoulli term: truncation below
// 1e-13 at z = 9.  Two logs and two divisions per argument, no data-dependent
// loop (no divergence across the wave).
struct LgPsi {
    double lg, psi;
};
__device__ __forceinline__ LgPsi d_lgamma_digamma(double x) {
    double p = 1.0, dp = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double t = x + (double)i;
        dp = dp * t + p;
        p *= t;
    }
    const double z = x + 8.0;
    const double iz = 1.0 / z, iz2 = iz * iz;
    const double lz = log(z);
    const double st =
        iz * (1.0 / 12.0 +
              iz2 * (-1.0 / 360.0 + iz2 * (1.0 / 1260.0 + iz2 * (-1.0 / 1680.0 + iz2 * (1.0 / 1188.0)))));
    const double ps =
        iz2 * (1.0 / 12.0 +
               iz2 * (-1.0 / 120.0 + iz2 * (1.0 / 252.0 + iz2 * (-1.0 / 240.0 + iz2 * (1.0 / 132.0)))));
    LgPsi r;
    r.lg = (z - 0.5) * lz - z + 0.91893853320467274178 + st - log(p);
    r.psi = lz - 0.5 * iz - ps - dp / p;
    return r;
}

// H(a, b) of Beta(a, b) (beta.rs:88-96):
// ln B(a, b) - (a-1) psi(a) - (b-1) psi(b) + (a+b-2) psi(a+b).
__device__ __forceinline__ double d_beta_entropy(double a, double b) {
    const double s = a + b;
    const LgPsi la = d_lgamma_digamma(a), lb = d_lgamma_digamma(b), ls = d_lgamma_digamma(s);
    return la.lg + lb.lg - ls.lg - (a - 1.0) * la.psi - (b - 1.0) * lb.psi + (s - 2.0) * ls.psi;
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
