/* Oracle (test infrastructure only): single-precision exp/ln used by the CPU
 * restatement of the splat render path.
 *
 * The reference evaluates exp/ln through its shader compiler's builtins
 * (src/project.rs:34, 105, 264-266, 313-316; src/raster.rs:187), whose last-ulp
 * behaviour is device dependent.  To make CPU and GPU results comparable bit
 * for bit, both sides evaluate one fixed algorithm: Cody-Waite range
 * reduction followed by the Cephes single-precision minimax polynomials
 * (S. L. Moshier, "Methods and Programs for Mathematical Functions", 1989;
 * Cephes expf.c / logf.c).  Polynomials are evaluated with explicit fmaf so
 * that no compiler contraction choice can change the bits.  Accuracy is pinned
 * against glibc (tests/test_oracle_math.py: <= 2 ulp over the ranges used).
 */
#ifndef SFO_MATH_H
#define SFO_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t sfo_f2b(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float sfo_b2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* 2^k for k in the normal exponent range [-126, 127]. */
static inline float sfo_pow2i(int k) { return sfo_b2f((uint32_t)(k + 127) << 23); }

static inline float sfo_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    /* n = round-half-up(x / ln2) */
    float n = floorf(x * 1.44269504088896341f + 0.5f);
    /* r = x - n ln2 in two pieces (ln2_hi has trailing zero bits). */
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
    if (k > 127) return (p * sfo_pow2i(127)) * sfo_pow2i(k - 127);
    if (k < -126) return (p * sfo_pow2i(k + 100)) * sfo_pow2i(-100);
    return p * sfo_pow2i(k);
}

static inline float sfo_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t u = sfo_f2b(x);
    if ((u >> 23) == 0) { /* subnormal: scale into the normal range */
        x = x * 33554432.0f; /* 2^25 */
        u = sfo_f2b(x);
        e = -25;
    }
    /* x = m * 2^(ex) with m in [0.5, 1) */
    e += (int)(u >> 23) - 126;
    float m = sfo_b2f((u & 0x007fffffu) | 0x3f000000u);
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

/* Rust `f32 as u32`: saturating, NaN -> 0. */
static inline uint32_t sfo_f2u(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

/* Rust f32::clamp (NaN passes through). */
static inline float sfo_clampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

#endif
