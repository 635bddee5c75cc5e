// Beta-distribution entropy in f64 on the device (src/seg/beta.rs:52-96):
// the reference evaluates it on the host to fill a 445x445 f32 table
// (beta.rs:100-144); here it is evaluated per splat in the EIG kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sfk {

// ln Gamma(x) and digamma(x) for x >= 1 in f64.  The reference evaluates
// them separately (Lanczos g = 7 and a recurrence to x >= 6 with a short
// asymptotic series, beta.rs:55-86) to fill its f32 table; here both come
// from one shift to z = x + 8 >= 9:
//   ln Gamma(x) = ln Gamma(z) - ln P,   psi(x) = psi(z) - P'/P,
//   P = prod_{i<8} (x + i),
// then Stirling's series for ln Gamma(z) and the asymptotic series for
// psi(z): truncation below 1e-13 at z = 9, no data-dependent loop.
// Shift-product terms of one argument x, normalized by z = x + 8 so they
// stay in range for any x >= 1: with t_i = (x + i) / z (i < 8),
// q = prod t_i = P / z^8 and dq = sum_i prod_{j != i} t_j = P' / z^7, so
// ln P = 8 ln z + ln q and P'/P = dq / (z q).
struct ShiftTerms {
    double q, dq;
};
__device__ __forceinline__ ShiftTerms d_shift_terms(double x, double iz) {
    ShiftTerms r{1.0, 0.0};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double t = (x + (double)i) * iz;
        r.dq = r.dq * t + r.q;
        r.q *= t;
    }
    return r;
}

// Stirling's series part of ln Gamma(z) past (z - 1/2) ln z - z + ln sqrt(2 pi),
// and the asymptotic series part of psi(z) past ln z - 1/(2z), to the 1/z^9
// (resp. 1/z^10) Bernoulli term.
__device__ __forceinline__ double d_stirling(double iz) {
    const double iz2 = iz * iz;
    return iz * (1.0 / 12.0 +
                 iz2 * (-1.0 / 360.0 + iz2 * (1.0 / 1260.0 + iz2 * (-1.0 / 1680.0 + iz2 * (1.0 / 1188.0)))));
}
__device__ __forceinline__ double d_psi_series(double iz) {
    const double iz2 = iz * iz;
    return iz2 * (1.0 / 12.0 +
                  iz2 * (-1.0 / 120.0 + iz2 * (1.0 / 252.0 + iz2 * (-1.0 / 240.0 + iz2 * (1.0 / 132.0)))));
}

// H(a, b) of Beta(a, b) (beta.rs:88-96):
// ln B(a, b) - (a-1) psi(a) - (b-1) psi(b) + (a+b-2) psi(a+b), from the
// shifted forms above for x in {a, b, s = a + b}, fused so
// that the three arguments share their divisions and shift logs:
//   ln Gamma(x) = (z - 17/2) ln z - z + ln sqrt(2 pi) + stirling(z) - ln q,
//   psi(x) = ln z - 1/(2z) - psi_series(z) - dq / (z q),
// with one division for the three 1/z, one log for ln(q_a q_b / q_s) (with
// its division), one division for the three dq / (z q) terms: four logs and
// three divisions instead of six and six.
__device__ __forceinline__ double d_beta_entropy(double a, double b) {
    const double s = a + b;
    const double za = a + 8.0, zb = b + 8.0, zs = s + 8.0;
    const double inv = 1.0 / (za * zb * zs);
    const double iza = zb * zs * inv, izb = za * zs * inv, izs = za * zb * inv;
    const ShiftTerms ta = d_shift_terms(a, iza), tb = d_shift_terms(b, izb),
                     ts = d_shift_terms(s, izs);
    const double la = log(za), lb = log(zb), ls = log(zs);
    const double lq = log(ta.q * tb.q / ts.q);
    constexpr double kHalfLn2Pi = 0.91893853320467274178;
    // ln B(a, b) = lnG(a) + lnG(b) - lnG(s); the three ln q terms are -lq.
    const double lnb = (za - 8.5) * la - za + d_stirling(iza) + (zb - 8.5) * lb - zb +
                       d_stirling(izb) - ((zs - 8.5) * ls - zs + d_stirling(izs)) + kHalfLn2Pi -
                       lq;
    const double wa = a - 1.0, wb = b - 1.0, ws = s - 2.0;
    // psi without its dq / (z q) term, then those three over one denominator:
    // - wa (dqa iza / qa) - wb (dqb izb / qb) + ws (dqs izs / qs).
    const double pa = la - 0.5 * iza - d_psi_series(iza);
    const double pb = lb - 0.5 * izb - d_psi_series(izb);
    const double ps = ls - 0.5 * izs - d_psi_series(izs);
    const double num = wa * ta.dq * iza * tb.q * ts.q + wb * tb.dq * izb * ta.q * ts.q -
                       ws * ts.dq * izs * ta.q * tb.q;
    const double rat = num / (ta.q * tb.q * ts.q);
    return lnb - wa * pa - wb * pb + ws * ps + rat;
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
