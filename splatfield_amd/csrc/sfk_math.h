// Single-precision exp/ln and conversions shared by every HIP kernel of the
// splat render path.
//
// The reference evaluates exp/ln with its shader compiler's builtins
// (src/project.rs:34, 105, 264-266, 313-316; src/raster.rs:187), whose last-ulp
// behaviour depends on the device.  Here one fixed algorithm is used so that
// results are reproducible on any host and device: Cody-Waite range reduction
// and the Cephes single-precision minimax polynomials (S. L. Moshier,
// "Methods and Programs for Mathematical Functions", 1989; Cephes expf/logf),
// evaluated with explicit fmaf.  The library is compiled with
// -ffp-contract=off, so every other a*b+c rounds twice, as written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SFK_HD __host__ __device__ __forceinline__

SFK_HD float sfk_pow2i(int k) { return __builtin_bit_cast(float, (uint32_t)(k + 127) << 23); }

SFK_HD float sfk_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    float n = floorf(x * 1.44269504088896341f + 0.5f);
    float r = fmaf(-n, 0.693359375f, x);
    r = fmaf(-n, -2.12194440e-4f, r);
    float z = r * r;
    float p = 1.9875691500e-4f;
    p = fmaf(p, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    p = fmaf(p, z, r);
    p = p + 1.0f;
    int k = (int)n;
    if (k > 127) return (p * sfk_pow2i(127)) * sfk_pow2i(k - 127);
    if (k < -126) return (p * sfk_pow2i(k + 100)) * sfk_pow2i(-100);
    return p * sfk_pow2i(k);
}

// exp(-p) for the rasterizers' Gaussian falloff, without sfk_expf's range
// branches.  For -p in [-87, 88] it is the same evaluation step for step
// (the scale 2^k stays a normal number, so it is built directly): identical
// bits.  NaN comes out as NaN.  -p above 88 is evaluated at 88: exp(88) =
// 1.65e38, so min(o * G, 0.999) = 0.999 exactly as with the true value for
// any opacity above 1e-35, and o * G >= 0.999 keeps every caller on the same
// side of its saturation test.  -p below -87 is evaluated at -87 (a tiny
// finite value): the callers evaluate whole waves but only use lanes whose
// power is at most 5.6, and a finite value keeps 0 * G at 0 elsewhere.
SFK_HD float sfk_expf_neg(float p) {
    const float a = -p;
    // Clamp without branches (min/max return the number for a NaN operand;
    // the NaN is restored by the final select).
    const float x = fminf(fmaxf(a, -87.0f), 88.0f);
    const float n = floorf(x * 1.44269504088896341f + 0.5f);
    float r = fmaf(-n, 0.693359375f, x);
    r = fmaf(-n, -2.12194440e-4f, r);
    const float z = r * r;
    float q = 1.9875691500e-4f;
    q = fmaf(q, r, 1.3981999507e-3f);
    q = fmaf(q, r, 8.3334519073e-3f);
    q = fmaf(q, r, 4.1665795894e-2f);
    q = fmaf(q, r, 1.6666665459e-1f);
    q = fmaf(q, r, 5.0000001201e-1f);
    q = fmaf(q, z, r);
    q = q + 1.0f;
    const float y = q * sfk_pow2i((int)n);
    return a == a ? y : a;
}

// exp(-p) for the lanes a responsibility rasterizer uses, whose G only
// enters alpha = min(o * G, 0.999) with o >= 1/255 (culled splats never
// reach a rasterizer), and only where the power is at most 5.6 or NaN:
//  * p <= 87: the same bits as sfk_expf(-p) (the scale 2^n is applied by
//    an exact ldexp, as sfk_expf's multiply by the normal 2^n is exact up to
//    the same final rounding);
//  * p < -88 and NaN p: exp(88) = 1.65e38, so o * G >= 0.999 and alpha is
//    0.999, as with the true value and as min(NaN, 0.999) gives for NaN;
//  * p > 87: unspecified (such lanes are discarded).
// Dropping sfk_expf_neg's lower clamp and NaN select saves their operations
// in the rasterizers' innermost loop.
SFK_HD float sfk_expf_neg_in(float p) {
    // min(-p, 88) written as -max(p, -88) (the same value, NaN -> 88), so
    // that the negation folds into the operands that use x.
    const float x = -fmaxf(p, -88.0f);
    const float n = floorf(x * 1.44269504088896341f + 0.5f);
    float r = fmaf(-n, 0.693359375f, x);
    r = fmaf(-n, -2.12194440e-4f, r);
    const float z = r * r;
    float q = 1.9875691500e-4f;
    q = fmaf(q, r, 1.3981999507e-3f);
    q = fmaf(q, r, 8.3334519073e-3f);
    q = fmaf(q, r, 4.1665795894e-2f);
    q = fmaf(q, r, 1.6666665459e-1f);
    q = fmaf(q, r, 5.0000001201e-1f);
    q = fmaf(q, z, r);
    q = q + 1.0f;
    return ldexpf(q, (int)n);
}

SFK_HD float sfk_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e = 0;
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u >> 23) == 0) {
        x = x * 33554432.0f;
        u = __builtin_bit_cast(uint32_t, x);
        e = -25;
    }
    e += (int)(u >> 23) - 126;
    float m = __builtin_bit_cast(float, (u & 0x007fffffu) | 0x3f000000u);
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    float fe = (float)e;
    float z = m * m;
    float y = 7.0376836292e-2f;
    y = fmaf(y, m, -1.1514610310e-1f);
    y = fmaf(y, m, 1.1676998740e-1f);
    y = fmaf(y, m, -1.2420140846e-1f);
    y = fmaf(y, m, 1.4249322787e-1f);
    y = fmaf(y, m, -1.6668057665e-1f);
    y = fmaf(y, m, 2.0000714765e-1f);
    y = fmaf(y, m, -2.4999993993e-1f);
    y = fmaf(y, m, 3.3333331174e-1f);
    y = (y * m) * z;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

// `f32 as u32` of the reference (Rust): saturating, NaN -> 0.  The
// conversion is only taken for 0 < v < 2^32, where it truncates exactly;
// written as selects, it compiles without branches (two compares, the
// hardware conversion, two selects).
SFK_HD uint32_t sfk_f2u(float v) {
    return v > 0.0f ? (v >= 4294967296.0f ? 0xffffffffu : (uint32_t)v) : 0u;
}

// f32::clamp of the reference: NaN passes through.
SFK_HD float sfk_clampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
