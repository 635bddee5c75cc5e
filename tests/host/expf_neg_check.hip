// Host-side check that the rasterizers' sfk_expf_neg(p) returns the same bits
// as sfk_expf(-p) (the oracle's exp) over every float p in [-88, 6] visited
// with the given stride, that NaN propagates, and that arguments past the
// clamp still saturate alpha = min(o * G, 0.999) for o >= 10/255; that the
// rasterizers' sfk_expf_neg_in has the same bits below p = 87 and gives
// alpha 0.999 for NaN; and that the branch-free sfk_f2u is Rust's
// saturating `f32 as u32`.
#include "../../splatfield_amd/csrc/sfk_math.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)strtoul(argv[1], nullptr, 10) : 1u;
    unsigned long long bad = 0, checked = 0;
    auto one = [&](uint32_t u) {
        float p;
        memcpy(&p, &u, 4);
        const float a = sfk_expf(-p), b = sfk_expf_neg(p), c = sfk_expf_neg_in(p);
        if (memcmp(&a, &b, 4) != 0 || memcmp(&a, &c, 4) != 0) {
            if (bad < 5) printf("p=%a exp=%a neg=%a neg_in=%a\n", p, a, b, c);
            ++bad;
        }
        ++checked;
    };
    for (uint64_t u = 0; u <= 0x40C00000u; u += stride) one((uint32_t)u);           // [0, 6]
    for (uint64_t u = 0x80000000u; u <= 0xC2B00000u; u += stride) one((uint32_t)u);  // [-88, 0)
    const float nan = __builtin_nanf("");
    const float r = sfk_expf_neg(nan);
    if (r == r) ++bad;
    // sfk_expf_neg_in(NaN) is exp(88): alpha = min(o * G, 0.999) is 0.999 for
    // every opacity a rasterized splat can have (o >= 1/255), as min(NaN,
    // 0.999) is.
    const float r_in = sfk_expf_neg_in(nan);
    const float e88 = sfk_expf(88.0f);
    if (memcmp(&r_in, &e88, 4) != 0) ++bad;
    if (!(fminf((1.0f / 255.0f) * r_in, 0.999f) == 0.999f)) ++bad;
    // sfk_expf_neg_in: the same bits as sfk_expf_neg below 87 and at the
    // upper clamp (p < -88) too.
    for (float p = -1e30f; p < 87.0f; p = p < -1.0f ? p / 1.7f : (p < 1.0f ? p + 0.37f : p * 1.3f)) {
        const float g = sfk_expf_neg(p), h = sfk_expf_neg_in(p);
        if (memcmp(&g, &h, 4) != 0) ++bad;
    }
    for (float p = -88.0f; p > -1e30f; p *= 1.7f)
        if (!(fminf((10.0f / 255.0f) * sfk_expf_neg(p), 0.999f) == 0.999f)) ++bad;
    // large powers: finite and below the contribution threshold
    for (float p = 87.0f; p < 1e30f; p *= 1.7f) {
        const float g = sfk_expf_neg(p);
        if (!(g >= 0.0f && g < 1.0f / 255.0f) || 0.0f * g != 0.0f) ++bad;
    }
    {
        const float g = sfk_expf_neg(__builtin_inff());
        if (!(g >= 0.0f && g < 1.0f / 255.0f)) ++bad;
    }
    // sfk_f2u (branch-free) against Rust's saturating `f32 as u32` spelled
    // with branches, on every float pattern visited with the same stride
    // plus the boundaries.
    auto f2u_ref = [](float v) -> uint32_t {
        if (!(v > 0.0f)) return 0u;
        if (v >= 4294967296.0f) return 0xffffffffu;
        return (uint32_t)v;
    };
    auto f2u_one = [&](uint32_t u) {
        float v;
        memcpy(&v, &u, 4);
        if (sfk_f2u(v) != f2u_ref(v)) {
            if (bad < 5) printf("f2u v=%a got=%u want=%u\n", v, sfk_f2u(v), f2u_ref(v));
            ++bad;
        }
        ++checked;
    };
    for (uint64_t u = 0; u <= 0xffffffffull; u += stride) f2u_one((uint32_t)u);
    const uint32_t edges[] = {0x00000000u, 0x80000000u, 0x3f800000u, 0x4f7fffffu, 0x4f800000u,
                              0x7f7fffffu, 0x7f800000u, 0xff800000u, 0x7fc00000u, 0x7f800001u,
                              0xffc00000u, 0xff800001u, 0x00000001u, 0xbf800000u};
    for (uint32_t u : edges) f2u_one(u);
    printf("checked %llu mismatches %llu\n", checked, bad);
    return bad != 0;
}
