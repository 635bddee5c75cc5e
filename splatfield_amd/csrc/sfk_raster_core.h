// Shared pieces of the forward and adjoint tile rasterizers: the pixel layout
// of a tile over its wavefronts and the per-wave splat lists.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfk_device.h"

namespace sfk {

constexpr float kInv255 = 1.0f / 255.0f;
// exp(-5.6) * 1.0 < 1/255: beyond this Gaussian power no splat (opacity <= 1)
// reaches the contribution threshold, whatever the last ulp of exp.
constexpr float kPowerSkip = 5.6f;

// Thread t of a TW x TW tile works on pixel (lx, ly): wavefront w covers the
// 8 x 8 block (w % (TW/8), w / (TW/8)), lane l its pixel (l % 8, l / 8).
// Square blocks let a wave skip every splat whose reach misses its block.
template <int TW>
__device__ __forceinline__ void tile_pixel(uint32_t t, uint32_t& lx, uint32_t& ly) {
    constexpr uint32_t BX = TW / 8;
    const uint32_t w = t / 64, l = t % 64;
    lx = (w % BX) * 8 + l % 8;
    ly = (w / BX) * 8 + l / 8;
}

// Axis-aligned box (x_lo, x_hi, y_lo, y_hi), in pixel coordinates, outside
// which a splat cannot reach the contribution threshold (raster.rs:183-190):
// alpha >= 1/255 needs o * exp(-power) >= 1/255, i.e. power <= ln(255 o), and
// the set {d : 0.5 d^T Q d <= tau}, Q = [[a, b], [b, c]] positive definite,
// has half-extents sqrt(2 tau c / det) and sqrt(2 tau a / det), det = ac - b^2.
// tau gets 0.05 of slack (5% in alpha) and the extents 0.1% + 0.01 px, far
// beyond the f32 rounding of the evaluated power for condition numbers below
// 1e4, so no pixel whose evaluated alpha reaches 1/255 lies outside the box.
// Any other case (Q not positive definite, ill-conditioned, NaN) gets an
// unbounded box: a NaN bound compares false and never culls.
// *tau_out is the same power bound tau for the exact ellipse test of
// wave_reach_list, +inf where that test does not apply (unbounded box).
__device__ __forceinline__ float4 reach_box(float4 geo, float4 aux, float* tau_out) {
    const float a = geo.z, b = geo.w, c = aux.x, o = aux.y;
    const float det = a * c - b * b;
    const float tr = a + c;
    *tau_out = INFINITY;
    if (!(a > 0.0f && det > 0.0f && det * 1e4f > tr * tr))
        return make_float4(-INFINITY, INFINITY, -INFINITY, INFINITY);
    const float tau = __logf(255.0f * o) + 0.05f;
    if (!(tau > 0.0f)) {
        // o < 1/255: alpha = o * G <= o never reaches the threshold (G <= 1
        // for a positive definite Q).
        return make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    }
    *tau_out = tau;
    const float hx = sqrtf(2.0f * tau * c / det) * 1.001f + 0.01f;
    const float hy = sqrtf(2.0f * tau * a / det) * 1.001f + 0.01f;
    return make_float4(geo.x - hx, geo.x + hx, geo.y - hy, geo.y + hy);
}

// Minimum of the power 0.5 (a x^2 + c y^2) + b x y over the rectangle
// [X0, X1] x [Y0, Y1] of mean-relative coordinates, for a positive definite
// conic: 0 when the rectangle holds the mean, else the least of the four
// edge minima (on an edge the power is a 1-D quadratic, minimized at its
// clamped vertex).  An inexact vertex only moves the evaluated point within
// the edge, by far less than the 0.05 of slack in tau.
__device__ __forceinline__ float power_min_rect(float X0, float X1, float Y0, float Y1, float a,
                                                float b, float c) {
    if (X0 <= 0.0f && X1 >= 0.0f && Y0 <= 0.0f && Y1 >= 0.0f) return 0.0f;
    const float ia = __builtin_amdgcn_rcpf(a), ic = __builtin_amdgcn_rcpf(c);
    auto at = [&](float x, float y) { return 0.5f * (a * x * x + c * y * y) + b * x * y; };
    const float yl = fminf(fmaxf(-b * X0 * ic, Y0), Y1), yr = fminf(fmaxf(-b * X1 * ic, Y0), Y1);
    const float xb = fminf(fmaxf(-b * Y0 * ia, X0), X1), xt = fminf(fmaxf(-b * Y1 * ia, X0), X1);
    return fminf(fminf(at(X0, yl), at(X1, yr)), fminf(at(xb, Y0), at(xt, Y1)));
}

// Per-wave list of the staged rows [0, cnt) that can reach the contribution
// threshold on this wave's 8 x 8 block of pixel centres [x0 + 0.5, x0 + 7.5]
// x [y0 + 0.5, y0 + 7.5], in staged (depth) order: the reach box meets the
// block and, where the ellipse test applies (s_tau finite), the least power
// over the block is within tau.  A splat failing either has alpha < 1/255 at
// every pixel of the block, so leaving it out changes nothing.  CH = 2 when
// conic a and c are staged halved.  Returns the list length.  Wave-local: the
// list is written and read by the same wavefront only.
template <int CH = 1>
__device__ __forceinline__ uint32_t wave_reach_list(const float4* s_box, const float* s_tau,
                                                    const float4* s_geo, const float4* s_aux,
                                                    uint32_t cnt, float x0, float y0,
                                                    uint16_t* list) {
    const uint32_t lane = lane_id();
    const float rx0 = x0 + 0.5f, rx1 = x0 + 7.5f, ry0 = y0 + 0.5f, ry1 = y0 + 7.5f;
    uint32_t n = 0;
    for (uint32_t base = 0; base < cnt; base += 64) {
        const uint32_t j = base + lane;
        bool hit = false;
        if (j < cnt) {
            float4 bx = s_box[j];
            hit = !(bx.y < rx0 || bx.x > rx1 || bx.w < ry0 || bx.z > ry1);
            const float tau = s_tau[j];
            if (hit && tau < INFINITY) {
                const float4 g = s_geo[j];
                const float c = (float)CH * s_aux[j].x;
                // (a NaN minimum does not cull, as a NaN box bound does not)
                hit = !(power_min_rect(rx0 - g.x, rx1 - g.x, ry0 - g.y, ry1 - g.y,
                                       (float)CH * g.z, g.w, c) > tau);
            }
        }
        const uint64_t m = __ballot(hit);
        if (hit) list[n + mask_rank(m)] = (uint16_t)j;
        n += (uint32_t)__popcll(m);
    }
    return n;
}

}  // namespace sfk
