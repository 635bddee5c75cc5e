// Beta-distribution entropy in f64 on the device (src/seg/beta.rs:52-98):
// the reference evaluates it on the host to fill a 445x445 f32 table
// (beta.rs:100-144); here it is evaluated per splat in the EIG kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sfk {

// Natural log of a positive, finite, normal double to about 1 ulp, in a third
// of the instructions of the library's log (which carries double-double
// arithmetic for the last bit): x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// f = m - 1 (exact), s = f / (2 + f), and
//   log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R)),
//   R = sum_{k>=1} 2 s^(2k) / (2k + 1)
// (2 atanh(s) = f - s f + s R, and s f = f^2/2 (1 - s) since f = s (2 + f)).
// R's Taylor series is cut after the s^18 term: |s| <= 0.1716 leaves the
// next term below 2^-55 of the result.  e ln 2 enters as e ln2_hi (ln2_hi
// has 32 significant bits, so the product is exact) plus e ln2_lo.
// Subnormal inputs are scaled by 2^54 first; zero gives -inf, +inf gives
// +inf, negative numbers and NaN give NaN.
__device__ __forceinline__ double d_log_pos(double x) {
    const bool sub = x < 2.2250738585072014e-308;
    int e;
    double m = frexp(sub ? x * 0x1p54 : x, &e);  // [0.5, 1)
    if (sub) e -= 54;
    if (m < 0.70710678118654757) {
        m += m;
        --e;
    }
    const double f = m - 1.0;
    const double d = 2.0 + f;
    // s = f / d: reciprocal estimate, two Newton steps, one residual correction.
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double s0 = f * r;
    const double sq = fma(r, fma(-s0, d, f), s0);
    const double z = sq * sq;
    double R = 2.0 / 19.0;
    R = fma(R, z, 2.0 / 17.0);
    R = fma(R, z, 2.0 / 15.0);
    R = fma(R, z, 2.0 / 13.0);
    R = fma(R, z, 2.0 / 11.0);
    R = fma(R, z, 2.0 / 9.0);
    R = fma(R, z, 2.0 / 7.0);
    R = fma(R, z, 2.0 / 5.0);
    R = fma(R, z, 2.0 / 3.0);
    R *= z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)e;
    constexpr double kLn2Hi = 0x1.62e42ff000000p-1;      // 0.6931471806019545
    constexpr double kLn2Lo = -4.2009150726810846e-11;   // ln 2 - kLn2Hi
    const double t = fma(sq, hfsq + R, dk * kLn2Lo);
    const double res = fma(dk, kLn2Hi, -((hfsq - t) - f));
    if (x > 0.0 && x <= 1.7976931348623157e308) return res;
    return x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan(""));
}

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
        r.dq = fma(r.dq, t, r.q);
        r.q *= t;
    }
    return r;
}

// Stirling's series part of ln Gamma(z) past (z - 1/2) ln z - z + ln sqrt(2 pi),
// and the asymptotic series part of psi(z) past ln z - 1/(2z), to the 1/z^9
// (resp. 1/z^10) Bernoulli term.
__device__ __forceinline__ double d_stirling(double iz) {
    const double iz2 = iz * iz;
    double p = 1.0 / 1188.0;
    p = fma(p, iz2, -1.0 / 1680.0);
    p = fma(p, iz2, 1.0 / 1260.0);
    p = fma(p, iz2, -1.0 / 360.0);
    p = fma(p, iz2, 1.0 / 12.0);
    return iz * p;
}
__device__ __forceinline__ double d_psi_series(double iz) {
    const double iz2 = iz * iz;
    double p = 1.0 / 132.0;
    p = fma(p, iz2, -1.0 / 240.0);
    p = fma(p, iz2, 1.0 / 252.0);
    p = fma(p, iz2, -1.0 / 120.0);
    p = fma(p, iz2, 1.0 / 12.0);
    return iz2 * p;
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
// a / b for positive, finite, normal b (and finite a), to about 1 ulp: the
// reciprocal estimate, two Newton steps and one residual correction (the
// library division adds scaling steps for extreme operands).
__device__ __forceinline__ double d_div_pos(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q0 = a * r;
    return fma(r, fma(-q0, b, a), q0);
}

__device__ __forceinline__ double d_beta_entropy(double a, double b) {
    const double s = a + b;
    const double za = a + 8.0, zb = b + 8.0, zs = s + 8.0;
    const double inv = d_div_pos(1.0, za * zb * zs);
    const double iza = zb * zs * inv, izb = za * zs * inv, izs = za * zb * inv;
    const ShiftTerms ta = d_shift_terms(a, iza), tb = d_shift_terms(b, izb),
                     ts = d_shift_terms(s, izs);
    const double la = d_log_pos(za), lb = d_log_pos(zb), ls = d_log_pos(zs);
    const double lq = d_log_pos(d_div_pos(ta.q * tb.q, ts.q));
    constexpr double kHalfLn2Pi = 0.91893853320467274178;
    // ln B(a, b) = lnG(a) + lnG(b) - lnG(s); the three ln q terms are -lq.
    const double ga = fma(za - 8.5, la, -za) + d_stirling(iza);
    const double gb = fma(zb - 8.5, lb, -zb) + d_stirling(izb);
    const double gs = fma(zs - 8.5, ls, -zs) + d_stirling(izs);
    const double lnb = ga + gb - gs + kHalfLn2Pi - lq;
    const double wa = a - 1.0, wb = b - 1.0, ws = s - 2.0;
    // psi without its dq / (z q) term, then those three over one denominator:
    // - wa (dqa iza / qa) - wb (dqb izb / qb) + ws (dqs izs / qs).
    const double pa = fma(-0.5, iza, la) - d_psi_series(iza);
    const double pb = fma(-0.5, izb, lb) - d_psi_series(izb);
    const double ps = fma(-0.5, izs, ls) - d_psi_series(izs);
    const double qab = ta.q * tb.q;
    const double num = fma(wa * ta.dq * iza, tb.q * ts.q,
                           fma(wb * tb.dq * izb, ta.q * ts.q, -(ws * ts.dq * izs * qab)));
    const double rat = d_div_pos(num, qab * ts.q);
    return fma(ws, ps, fma(-wb, pb, fma(-wa, pa, lnb))) + rat;
}

// max(0, H0 - H(a k, b k)), k = 1 + eps/(a+b), eps = f32(q)/scale; 0 when
// eps = 0 (beta.rs:172-186).  H0 = H(a, b).
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
