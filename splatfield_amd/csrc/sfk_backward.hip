// Backward (adjoint) pass of the splat render for gfx950.  The reference is
// forward-only (docs/00_pipeline.md:1); this differentiates its forward
// function — project_splats (src/project.rs:234-388) and the front-to-back
// compositor (src/raster.rs:138-229) — with respect to means, log-scales,
// quaternions (raw, as the kernel uses them), opacity logits and SH.
//
// Raster adjoint: one workgroup per tile walks the tile's isects back to
// front in LDS-staged chunks.  Per-pixel transmittance is recomputed from the
// forward's final T (T_k = T_{k+1} / (1 - alpha_k)), never stored per isect;
// the forward keeps only final T and the index past the last contributor per
// pixel.  Per-splat screen-space gradients (mean2d, conic, opacity, rgb) are
// summed across the wavefront, then across the workgroup in LDS, and added to
// global memory once per (splat, tile).
#include "sfk_device.h"
#include "sfk_internal.h"
#include "sfk_math.h"
#include "sfk_project_core.h"
#include "sfk_raster_core.h"

namespace sfk {

namespace {

__global__ __launch_bounds__(256) void k_loss_l2(const float4* __restrict__ img,
                                                 const float4* __restrict__ target, uint32_t np,
                                                 float4* __restrict__ d_img,
                                                 double* __restrict__ partials) {
    __shared__ double s_w[4];
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    double sq = 0.0;
    if (p < np) {
        float4 a = img[p], b = target[p];
        float4 d = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
        d_img[p] = make_float4(2.0f * d.x, 2.0f * d.y, 2.0f * d.z, 2.0f * d.w);
        sq = (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
    }
    sq = wave_sum_f64(sq);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sq;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

__global__ __launch_bounds__(256) void k_sum_f64(const double* __restrict__ partials, uint32_t np,
                                                 double* __restrict__ out) {
    __shared__ double s_w[4];
    double sum = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += 256) sum += partials[i];
    sum = wave_sum_f64(sum);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) *out = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

constexpr int kBwdMembers = 4;

struct BwdArgs {
    uint32_t w, h, tiles_x, tiles_y, n;
    const uint32_t* tile_starts;
    const uint32_t* tile_order;  // optional launch order of the tiles (heaviest first)
    const uint32_t* gids;
    const float* proj;
    const float* final_t;
    const uint32_t* n_contrib;
    const float* d_img;
    float* grad2d;  // 9 planes of n: mx, my, conic a, conic b, conic c, opacity, r, g, b
    // Fused L2 loss (d_img == nullptr): dL/dC = 2 (img - target) per pixel
    // and one f64 loss partial per tile.
    const float* img;
    const float* target;
    double* loss_partials;  // one per tile
};

// Fixed-order f64 sum of one value per thread over a workgroup of NT threads.
template <int NT>
__device__ __forceinline__ double block_sum_f64(double v, double* s_w) {
    v = wave_sum_f64(v);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t += s_w[w];
    __syncthreads();
    return t;
}

template <int M>
struct MemberSum;
template <>
struct MemberSum<8> {
    __device__ static float sum(const float (&v)[8]) { return wave_sum8_transposed(v); }
    static constexpr uint32_t kShift = 3;  // lane l holds member l >> 3
};
template <>
struct MemberSum<4> {
    __device__ static float sum(const float (&v)[4]) { return wave_sum4_transposed(v); }
    static constexpr uint32_t kShift = 4;  // lane l holds member l >> 4
};

// M = members per group: the gradients of M splats are summed over the wave
// in one transposed pass; M = 4 halves the registers held across a group.
template <int TW, int M>
__global__ __launch_bounds__(TW * TW) void k_raster_bwd(BwdArgs a) {
    constexpr int NT = TW * TW;
    constexpr int NW = NT / 64;
    __shared__ float4 s_geo[NT];
    __shared__ float4 s_aux[NT];
    __shared__ float4 s_box[NT];
    __shared__ float s_tau[NT];
    __shared__ uint16_t s_list[NW][NT];
    __shared__ float s_blue[NT];
    __shared__ uint32_t s_gid[NT];
    __shared__ float s_acc[9][NT];
    __shared__ uint32_t s_last;
    __shared__ double s_w[NW];

    const uint32_t ntiles = a.tiles_x * a.tiles_y;
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, ntiles);
    const uint32_t tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const uint32_t t = threadIdx.x;
    uint32_t lx, ly;
    tile_pixel<TW>(t, lx, ly);
    const uint32_t px = tx * TW + lx, py = ty * TW + ly;
    const bool in_bounds = px < a.w && py < a.h;
    const uint32_t start = a.tile_starts[tile];
    const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
    const float bx0 = (float)(tx * TW + (lx & ~7u)), by0 = (float)(ty * TW + (ly & ~7u));
    uint16_t* list = s_list[t / 64];

    uint32_t last = start;
    float T = 1.0f, Tfin = 1.0f;
    float4 dC = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in_bounds) {
        size_t p = (size_t)py * a.w + px;
        last = a.n_contrib[p];
        Tfin = a.final_t[p];
        T = Tfin;
        if (a.d_img) dC = reinterpret_cast<const float4*>(a.d_img)[p];
    }
    if (!a.d_img) {
        // Same arithmetic as the standalone loss: d = img - target, dC = 2 d.
        double sq = 0.0;
        if (in_bounds) {
            const size_t p = (size_t)py * a.w + px;
            const float4 im = reinterpret_cast<const float4*>(a.img)[p];
            const float4 tg = reinterpret_cast<const float4*>(a.target)[p];
            const float4 d = make_float4(im.x - tg.x, im.y - tg.y, im.z - tg.z, im.w - tg.w);
            dC = make_float4(2.0f * d.x, 2.0f * d.y, 2.0f * d.z, 2.0f * d.w);
            sq = (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
        }
        // One partial per tile; the projection adjoint, which runs next,
        // adds them up in tile order.
        const double part = block_sum_f64<NT>(sq, s_w);
        if (t == 0) a.loss_partials[tile] = part;
    }
    if (t == 0) s_last = start;
    __syncthreads();
    atomicMax(&s_last, last);
    __syncthreads();
    const uint32_t block_last = s_last;
    // Past its own last contributor a wave has nothing to do.
    uint32_t wave_last = last;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wave_last = max(wave_last, (uint32_t)__shfl_xor(wave_last, o, 64));
    // Equal in every lane: made uniform, the list length and the member loop
    // below are scalar (no per-lane loop control or list addressing).
    wave_last = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_last);
    float acc_r = 0.f, acc_g = 0.f, acc_b = 0.f;
    // d loss / d alpha through the final transmittance: Tfin dC.w / (1 - alpha).
    const float tfin_da = Tfin * dC.w;
    const float4* proj = reinterpret_cast<const float4*>(a.proj);

    // The splat ids of the next chunk are loaded while the current one is
    // walked (one memory round trip less at each staging).
    auto chunk_begin = [&](uint32_t ce) { return ce - start > (uint32_t)NT ? ce - NT : start; };
    // Loads at clamped indices, with no branch around them: the wait
    // counters then see the same loads on every path and need not drain.
    uint32_t g_next = 0;
    if (block_last > start) {
        const uint32_t cb0 = chunk_begin(block_last);
        g_next = a.gids[cb0 + min(t, block_last - cb0 - 1u)];
    }
    for (uint32_t cend = block_last; cend > start;) {
        const uint32_t cbeg = chunk_begin(cend);
        const uint32_t cnt = cend - cbeg;
        const uint32_t g = g_next;
        lds_barrier();
        const float4* row = proj + (size_t)g * 3;  // lanes past cnt: a copy of a live id
        const float4 geo = row[0], aux = row[1];
        const float blue = row[2].x;
        // Issued after the rows (the scheduling barrier keeps it there): loads
        // complete in issue order, so this one stays in flight past the wait
        // for the rows.
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t nb = chunk_begin(cbeg);
        g_next = a.gids[cbeg > start ? nb + min(t, cbeg - nb - 1u) : cbeg];
        if (t < cnt) {
            s_box[t] = reach_box(geo, aux, &s_tau[t]);
            // Conic a and c are staged halved: the power below is then
            // dx (a/2 dx + b dy) + c/2 dy^2, five operations.
            s_geo[t] = make_float4(geo.x, geo.y, 0.5f * geo.z, geo.w);
            s_aux[t] = make_float4(0.5f * aux.x, aux.y, aux.z, aux.w);
            s_blue[t] = blue;
            s_gid[t] = g;
#pragma unroll
            for (int k = 0; k < 9; ++k) s_acc[k][t] = 0.0f;
        }
        __syncthreads();
        const uint32_t lim = wave_last > cbeg ? min(cnt, wave_last - cbeg) : 0u;
        // This lane's contributors in the chunk: staged rows j < mine.
        const uint32_t mine = last > cbeg ? last - cbeg : 0u;
        const uint32_t nl = wave_reach_list<2>(s_box, s_tau, s_geo, s_aux, lim, bx0, by0, list);
        // Listed splats in groups of 8, back to front: member s of the group
        // whose top is `itop` is list entry itop - 1 - s.  The 9 x 8 per-pixel
        // terms are reduced over the wave in one transposed pass; lane l ends
        // with the sums of member l >> 3.
        for (uint32_t itop = nl; itop > 0; itop = itop > (uint32_t)M ? itop - M : 0) {
            float g[9][M];
            bool has = false;
#pragma unroll
            for (int s = 0; s < M; ++s) {
                // Member s, or (masked off) the group's top entry again.
                const bool member = s < (int)itop;
                const uint32_t j = list[member ? itop - 1 - s : itop - 1];
                const float4 geo = s_geo[j];
                const float4 aux = s_aux[j];
                const float dx = geo.x - pxf, dy = geo.y - pyf;
                // The forward's power up to rounding (fused here); a rounding
                // difference can only move the kPowerSkip test where alpha
                // is below 1/255 anyway.
                const float power = fmaf(dx, fmaf(geo.z, dx, geo.w * dy), aux.x * dy * dy);
                const bool in = member && j < mine && !(power > kPowerSkip);
                // Straight-line evaluation for the whole wave, every member
                // (a wave-wide skip test costs more than it saves); lanes
                // that do not contribute keep their state through selects.
                // exp(-power) by the hardware exp2: within an ulp or two of the
                // forward's exp, so alpha and T agree with the forward to
                // rounding (a contributor test can only differ where alpha is
                // within an ulp of 1/255).  The argument is capped so that a
                // NaN or negative power gives a finite G (alpha then clamps
                // to 0.999 as in the forward and dm = 0 keeps dm * G at 0).
                const float G = __builtin_amdgcn_exp2f(fminf(-power * 1.44269504f, 126.0f));
                const float araw = aux.y * G;
                const float alpha = fminf(araw, 0.999f);
                const bool con = in && alpha >= kInv255;
                has |= con;
                const float om = 1.0f - alpha;  // in [0.001, 1]
                const float rom = __builtin_amdgcn_rcpf(om);
                const float Tn = T * rom;
                T = con ? Tn : T;
                const float cb = s_blue[j];
                const float wr = aux.z - acc_r, wg = aux.w - acc_g, wb = cb - acc_b;
                // Colour weight: 0 on lanes that do not contribute.
                const float wgt = con ? alpha * Tn : 0.0f;
                g[6][s] = wgt * dC.x;
                g[7][s] = wgt * dC.y;
                g[8][s] = wgt * dC.z;
                // Gradient-only arithmetic (alpha, T and the contributor
                // tests above keep the forward's exact operation sequence)
                // uses fused multiply-adds.
                const float dal =
                    fmaf(tfin_da, rom, Tn * fmaf(wb, dC.z, fmaf(wg, dC.y, wr * dC.x)));
                // Back-to-front colour accumulators: acc' = acc + alpha (c - acc).
                acc_r = con ? fmaf(alpha, wr, acc_r) : acc_r;
                acc_g = con ? fmaf(alpha, wg, acc_g) : acc_g;
                acc_b = con ? fmaf(alpha, wb, acc_b) : acc_b;
                // d alpha / d(opacity, power) are zero where the 0.999 clamp
                // binds; the power terms are reduced as moments of dpow in
                // (dx, dy) and turned into mean / conic gradients per member
                // after the reduction:
                //   d mean = (a Sx + b Sy, c Sy + b Sx), d conic = (Sxx/2, Sxy, Syy/2).
                const float dm = (con && araw < 0.999f) ? dal : 0.0f;
                const float dpow = -dm * alpha;
                const float px_ = dpow * dx, py_ = dpow * dy;
                g[5][s] = dm * G;
                g[0][s] = px_;
                g[1][s] = py_;
                g[2][s] = px_ * dx;
                g[3][s] = px_ * dy;
                g[4][s] = py_ * dy;
            }
            if (__ballot(has)) {
                const uint32_t lane = lane_id();
                const uint32_t ms = lane >> MemberSum<M>::kShift;
                const bool writer = (lane & ((1u << MemberSum<M>::kShift) - 1u)) == 0 && ms < itop;
                const uint32_t j = writer ? list[itop - 1 - ms] : 0u;
                float r[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    r[k] = MemberSum<M>::sum(g[k]);
                    // Complete the sum here: sunk into the writers' branch its
                    // last DPP step would no longer fold into the add.
                    asm volatile("" : "+v"(r[k]));
                }
                // Every lane forms the member gradients (same cost per wave as
                // the writers alone); writers add them without testing for
                // zeros, which would cost a compare and branch per plane.
                const float4 geo = s_geo[j];
                const float qc = 2.0f * s_aux[j].x;  // staged halved
                float out[9];
                out[0] = fmaf(2.0f * geo.z, r[0], geo.w * r[1]);  // d mx
                out[1] = fmaf(qc, r[1], geo.w * r[0]);            // d my
                out[2] = 0.5f * r[2];                      // d conic a
                out[3] = r[3];                             // d conic b
                out[4] = 0.5f * r[4];                      // d conic c
#pragma unroll
                for (int k = 5; k < 9; ++k) out[k] = r[k];
                if (writer) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) atomicAdd(&s_acc[k][j], out[k]);
                }
            }
        }
        __syncthreads();
        if (t < cnt) {
            uint32_t gid = s_gid[t];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                float v = s_acc[k][t];
                if (v != 0.0f) unsafeAtomicAdd(&a.grad2d[(size_t)k * a.n + gid], v);
            }
        }
        cend = cbeg;
    }
}

constexpr float kShC0 = 0.2820948f;
constexpr float kShC1 = 0.48860252f;
constexpr float kShC20 = 0.9461747f;
constexpr float kShC21 = 0.31539157f;
constexpr float kShC22 = -1.0925485f;
constexpr float kShC23 = 0.54627424f;
constexpr float kK9 = -0.5900436f;

// Projection adjoint, one thread per splat.  CHS = SH coefficients per channel.
template <int CHS>
__global__ __launch_bounds__(256) void k_project_bwd(sfk_view v, const float* __restrict__ A,
                                                     const float* __restrict__ shs,
                                                     const float* __restrict__ dcol, uint32_t n,
                                                     float* __restrict__ g2,
                                                     float* __restrict__ gA,
                                                     float* __restrict__ gS,
                                                     const double* __restrict__ loss_partials,
                                                     uint32_t loss_np,
                                                     double* __restrict__ loss_out) {
    if (loss_out && blockIdx.x == 0) {
        // Fused L2 loss: fixed-order total of the adjoint raster's tile
        // partials (complete at this kernel's start).
        __shared__ double s_w[4];
        double sum = 0.0;
        for (uint32_t k = threadIdx.x; k < loss_np; k += 256) sum += loss_partials[k];
        sum = block_sum_f64<256>(sum, s_w);
        if (threadIdx.x == 0) *loss_out = sum;
    }
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    // Every per-splat input is loaded before any is used, and the loaded
    // values are pinned ahead of the visibility branch: left to itself the
    // compiler sinks the loads that only the visible path needs below that
    // branch, where each group waits out its own memory round trip.
    float ar[11], G2[9], J[9];
#pragma unroll
    for (int k = 0; k < 11; ++k) ar[k] = A[(size_t)k * N + i];
#pragma unroll
    for (int k = 0; k < 9; ++k) G2[k] = g2[(size_t)k * N + i];
    if (dcol) {
#pragma unroll
        for (int k = 0; k < 9; ++k) J[k] = dcol[(size_t)k * N + i];
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) J[k] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 11; ++k) asm volatile("" : "+v"(ar[k]));
#pragma unroll
    for (int k = 0; k < 9; ++k) asm volatile("" : "+v"(G2[k]));
#pragma unroll
    for (int k = 0; k < 9; ++k) asm volatile("" : "+v"(J[k]));
    float mx = ar[0], my = ar[1], mz = ar[2];
    float qw = ar[3], qx = ar[4], qy = ar[5], qz = ar[6];
    float s[3] = {sfk_expf(ar[7]), sfk_expf(ar[8]), sfk_expf(ar[9])};
    float op = 1.0f / (1.0f + sfk_expf(-ar[10]));
    float cam[3];
    cam[0] = v.rot0[0] * mx + v.rot0[1] * my + v.rot0[2] * mz + v.trans[0];
    cam[1] = v.rot1[0] * mx + v.rot1[1] * my + v.rot1[2] * mz + v.trans[1];
    cam[2] = v.rot2[0] * mx + v.rot2[1] * my + v.rot2[2] * mz + v.trans[2];
    bool vis = !(op < 10.0f / 255.0f) && !(cam[2] <= 0.1f);
    if (!vis) {
        for (int k = 0; k < 11; ++k) gA[(size_t)k * N + i] = 0.0f;
        for (int k = 0; k < 3 * CHS; ++k) gS[(size_t)k * N + i] = 0.0f;
        return;
    }
    // Leave zeros behind in the screen-space gradients just read: the next
    // backward accumulates into a cleared buffer without a fill pass.
#pragma unroll
    for (int k = 0; k < 9; ++k) g2[(size_t)k * N + i] = 0.0f;

    const float iz = 1.0f / cam[2];
    float x2 = qx * qx, y2 = qy * qy, z2 = qz * qz, xy = qx * qy, xz = qx * qz, yz = qy * qz;
    float wx = qw * qx, wy = qw * qy, wz = qw * qz;
    float R[3][3] = {{1.0f - 2.0f * (y2 + z2), 2.0f * (xy - wz), 2.0f * (xz + wy)},
                     {2.0f * (xy + wz), 1.0f - 2.0f * (x2 + z2), 2.0f * (yz - wx)},
                     {2.0f * (xz - wy), 2.0f * (yz + wx), 1.0f - 2.0f * (x2 + y2)}};
    float M[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) M[r][k] = R[r][k] * s[k];
    const float* r0 = v.rot0;
    const float* r1 = v.rot1;
    const float* r2 = v.rot2;
    float limx = 1.3f * v.img[0] / (2.0f * v.focal[0]);
    float limy = 1.3f * v.img[1] / (2.0f * v.focal[1]);
    float u = cam[0] * iz, vv = cam[1] * iz;
    bool uc = u < -limx || u > limx, vc = vv < -limy || vv > limy;
    u = sfk_clampf(u, -limx, limx);
    vv = sfk_clampf(vv, -limy, limy);
    float fxz = v.focal[0] * iz, fyz = v.focal[1] * iz;
    float t0[3], t1[3], j0[3], j1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t0[k] = fxz * r0[k] - fxz * u * r2[k];
        t1[k] = fyz * r1[k] - fyz * vv * r2[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        j0[k] = t0[0] * M[0][k] + t0[1] * M[1][k] + t0[2] * M[2][k];
        j1[k] = t1[0] * M[0][k] + t1[1] * M[1][k] + t1[2] * M[2][k];
    }
    float ca = j0[0] * j0[0] + j0[1] * j0[1] + j0[2] * j0[2] + 0.3f;
    float cb = j0[0] * j1[0] + j0[1] * j1[1] + j0[2] * j1[2];
    float cc = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2] + 0.3f;
    float detr = ca * cc - cb * cb;
    bool dclamp = detr < 1e-6f;
    float det = dclamp ? 1e-6f : detr;

    // colour -> SH, view direction
    float d[3] = {mx - v.world_pos[0], my - v.world_pos[1], mz - v.world_pos[2]};
    float vlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float ivl = 1.0f / vlen;
    float x = d[0] * ivl, y = d[1] * ivl, z = d[2] * ivl;
    float B[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) B[k] = 0.f;
    B[0] = kShC0;
    if constexpr (CHS >= 4) {
        B[1] = -kShC1 * y; B[2] = kShC1 * z; B[3] = -kShC1 * x;
    }
    if constexpr (CHS >= 9) {
        B[4] = kShC23 * 2.0f * x * y; B[5] = kShC22 * z * y; B[6] = kShC20 * z * z - kShC21;
        B[7] = kShC22 * z * x; B[8] = kShC23 * (x * x - y * y);
    }
    if constexpr (CHS >= 16) {
        float c1x = x * x - y * y, c1y = 2.f * x * y;
        float tt0 = -2.285229f * z * z + 0.4570458f, tt1 = 1.4453057f * z;
        B[9] = kK9 * x * c1y + y * c1x;
        B[10] = kK9 * x * c1x - y * c1y;
        B[11] = tt0 * y; B[12] = tt0 * x;
        B[13] = z * (1.8658817f * z * z - 1.119529f);
        B[14] = tt1 * c1y; B[15] = tt1 * c1x;
    }
#pragma unroll
    for (int k = 0; k < CHS; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) gS[(size_t)(3 * k + c) * N + i] = B[k] * G2[6 + c];
    // Gradient w.r.t. the unit view direction through the colour: from the
    // forward's d colour / d direction when it kept it (9 floats read
    // instead of the 3 x CHS SH coefficients), else from the coefficients.
    float gd0 = 0.f, gd1 = 0.f, gd2 = 0.f;
    if (dcol) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float gc = G2[6 + c];
            gd0 += gc * J[3 * c + 0];
            gd1 += gc * J[3 * c + 1];
            gd2 += gc * J[3 * c + 2];
        }
    } else {
        float shc[3 * CHS];
#pragma unroll
        for (int k = 0; k < 3 * CHS; ++k) shc[k] = shs[(size_t)k * N + i];
        float D[16][3];
        sh_basis_dir_grad<CHS>(x, y, z, D);
#pragma unroll
        for (int k = 0; k < CHS; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float sv = shc[3 * k + c] * G2[6 + c];
                gd0 += sv * D[k][0];
                gd1 += sv * D[k][1];
                gd2 += sv * D[k][2];
            }
    }
    float dd = gd0 * x + gd1 * y + gd2 * z;
    float gm[3] = {(gd0 - x * dd) * ivl, (gd1 - y * dd) * ivl, (gd2 - z * dd) * ivl};

    gA[10 * N + i] = G2[5] * op * (1.0f - op);

    float gx = G2[2], gy = G2[3], gz = G2[4];
    float inv = 1.0f / det, inv2 = inv * inv;
    float ga, gb, gc;
    if (!dclamp) {
        ga = gx * (-cc * cc * inv2) + gy * (cb * cc * inv2) + gz * (inv - ca * cc * inv2);
        gb = gx * (2.f * cb * cc * inv2) + gy * (-inv - 2.f * cb * cb * inv2) + gz * (2.f * ca * cb * inv2);
        gc = gx * (inv - ca * cc * inv2) + gy * (ca * cb * inv2) + gz * (-ca * ca * inv2);
    } else {
        ga = gz * inv; gb = -gy * inv; gc = gx * inv;
    }
    float gj0[3], gj1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gj0[k] = 2.f * ga * j0[k] + gb * j1[k];
        gj1[k] = gb * j0[k] + 2.f * gc * j1[k];
    }
    float gt0[3] = {0.f, 0.f, 0.f}, gt1[3] = {0.f, 0.f, 0.f}, gR[3][3];
    float gs[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float gM = t0[r] * gj0[k] + t1[r] * gj1[k];
            gt0[r] += M[r][k] * gj0[k];
            gt1[r] += M[r][k] * gj1[k];
            gs[k] += R[r][k] * gM;
            gR[r][k] = gM * s[k];
        }
#pragma unroll
    for (int k = 0; k < 3; ++k) gA[(size_t)(7 + k) * N + i] = gs[k] * s[k];
    float gqw = 0.f, gqx = 0.f, gqy = 0.f, gqz = 0.f;
    gqy += gR[0][0] * (-4.f * qy); gqz += gR[0][0] * (-4.f * qz);
    gqx += gR[0][1] * 2.f * qy; gqy += gR[0][1] * 2.f * qx; gqw -= gR[0][1] * 2.f * qz; gqz -= gR[0][1] * 2.f * qw;
    gqx += gR[0][2] * 2.f * qz; gqz += gR[0][2] * 2.f * qx; gqw += gR[0][2] * 2.f * qy; gqy += gR[0][2] * 2.f * qw;
    gqx += gR[1][0] * 2.f * qy; gqy += gR[1][0] * 2.f * qx; gqw += gR[1][0] * 2.f * qz; gqz += gR[1][0] * 2.f * qw;
    gqx += gR[1][1] * (-4.f * qx); gqz += gR[1][1] * (-4.f * qz);
    gqy += gR[1][2] * 2.f * qz; gqz += gR[1][2] * 2.f * qy; gqw -= gR[1][2] * 2.f * qx; gqx -= gR[1][2] * 2.f * qw;
    gqx += gR[2][0] * 2.f * qz; gqz += gR[2][0] * 2.f * qx; gqw -= gR[2][0] * 2.f * qy; gqy -= gR[2][0] * 2.f * qw;
    gqy += gR[2][1] * 2.f * qz; gqz += gR[2][1] * 2.f * qy; gqw += gR[2][1] * 2.f * qx; gqx += gR[2][1] * 2.f * qw;
    gqx += gR[2][2] * (-4.f * qx); gqy += gR[2][2] * (-4.f * qy);
    gA[3 * N + i] = gqw; gA[4 * N + i] = gqx; gA[5 * N + i] = gqy; gA[6 * N + i] = gqz;

    float fx = v.focal[0], fy = v.focal[1];
    float giz = 0.f, gu = 0.f, gv = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        giz += gt0[k] * fx * (r0[k] - u * r2[k]) + gt1[k] * fy * (r1[k] - vv * r2[k]);
        gu += gt0[k] * (-fxz * r2[k]);
        gv += gt1[k] * (-fyz * r2[k]);
    }
    float gcam[3] = {0.f, 0.f, 0.f};
    if (!uc) { gcam[0] += gu * iz; giz += gu * cam[0]; }
    if (!vc) { gcam[1] += gv * iz; giz += gv * cam[1]; }
    gcam[0] += G2[0] * fx * iz; giz += G2[0] * fx * cam[0];
    gcam[1] += G2[1] * fy * iz; giz += G2[1] * fy * cam[1];
    gcam[2] += giz * (-iz * iz);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        gA[(size_t)k * N + i] = gm[k] + r0[k] * gcam[0] + r1[k] * gcam[1] + r2[k] * gcam[2];
}

}  // namespace

hipError_t launch_loss_l2(hipStream_t s, const float* img, const float* target, uint32_t np,
                          float* d_img, double* partials, double* out) {
    uint32_t nb = div_up(np ? np : 1, 256);
    if (np) k_loss_l2<<<nb, 256, 0, s>>>(reinterpret_cast<const float4*>(img),
                                         reinterpret_cast<const float4*>(target), np,
                                         reinterpret_cast<float4*>(d_img), partials);
    else nb = 0;
    if (out) k_sum_f64<<<1, 256, 0, s>>>(partials, nb, out);
    return hipGetLastError();
}

uint32_t loss_num_partials(uint32_t np) { return div_up(np ? np : 1, 256); }

hipError_t launch_raster_bwd(hipStream_t s, const RasterArgs& ra, const float* final_t,
                             const uint32_t* n_contrib, const float* d_img, uint32_t n,
                             float* grad2d, const BwdLoss* loss) {
    BwdArgs a;
    a.w = ra.w; a.h = ra.h; a.tiles_x = ra.tiles_x; a.tiles_y = ra.tiles_y; a.n = n;
    a.tile_starts = ra.tile_starts; a.tile_order = ra.tile_order; a.gids = ra.gids; a.proj = ra.proj;
    a.final_t = final_t; a.n_contrib = n_contrib; a.d_img = d_img; a.grad2d = grad2d;
    a.img = a.target = nullptr;
    a.loss_partials = nullptr;
    if (!d_img) {
        if (!loss) return hipErrorInvalidValue;
        a.img = loss->img; a.target = loss->target; a.loss_partials = loss->partials;
    }
    uint32_t nt = ra.tiles_x * ra.tiles_y;
    if (nt == 0) return hipSuccess;
    constexpr int M = kBwdMembers;
    switch (ra.tile_w) {
        case 8: k_raster_bwd<8, M><<<nt, 64, 0, s>>>(a); break;
        case 16: k_raster_bwd<16, M><<<nt, 256, 0, s>>>(a); break;
        case 32: k_raster_bwd<32, M><<<nt, 1024, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_project_bwd(hipStream_t s, const sfk_view& v, const float* attrs,
                              const float* sh, uint32_t n, uint32_t sh_per_ch,
                              float* grad2d, float* g_attrs, float* g_sh,
                              const BwdLoss* loss, const float* dcol) {
    if (n == 0) return hipSuccess;
    dim3 grid(div_up(n, 256));
    const double* lp = loss ? loss->partials : nullptr;
    const uint32_t lnp = loss ? loss->num_partials : 0u;
    double* lo = loss ? loss->out : nullptr;
    switch (sh_per_ch) {
        case 1: k_project_bwd<1><<<grid, 256, 0, s>>>(v, attrs, sh, dcol, n, grad2d, g_attrs, g_sh, lp, lnp, lo); break;
        case 4: k_project_bwd<4><<<grid, 256, 0, s>>>(v, attrs, sh, dcol, n, grad2d, g_attrs, g_sh, lp, lnp, lo); break;
        case 9: k_project_bwd<9><<<grid, 256, 0, s>>>(v, attrs, sh, dcol, n, grad2d, g_attrs, g_sh, lp, lnp, lo); break;
        case 16: k_project_bwd<16><<<grid, 256, 0, s>>>(v, attrs, sh, dcol, n, grad2d, g_attrs, g_sh, lp, lnp, lo); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sfk
