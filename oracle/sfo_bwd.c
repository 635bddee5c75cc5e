/* Oracle (test infrastructure only): f64 loss of the splat render and its
 * hand-derived adjoint, for checking the GPU backward pass.
 *
 * The reference has no backward pass (docs/00_pipeline.md:1); the forward
 * function being differentiated is the reference's: project_splats
 * (src/project.rs:234-388), the compositor (src/raster.rs:138-229) and an L2
 * loss over the f32 RGBA image (r, g, b, 1 - T).  The function is piecewise
 * smooth: culls, the u/v clamp (project.rs:192-195), the det clamp
 * (project.rs:225), alpha >= 1/255, the 0.999 alpha cap and the T < 1/255
 * stop are branch decisions.  sfo_bwd_loss(record = 1) evaluates in f64 and
 * records every decision; record = 0 replays them, so central finite
 * differences measure the derivative inside the recorded piece — the
 * derivative the adjoint computes.  The tile lists (which splats a tile
 * visits, in which order) come from the f32 prepare and stay fixed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sfo.h"
#include "sfo_math.h"

#include <pthread.h>

typedef struct {
    double mx, my, ca, cb, cc;   /* mean2d, cov2d */
    double qa, qb, qc;           /* conic */
    double op, r, g, b;          /* opacity, colour (+0.5) */
    /* intermediates for the adjoint */
    double cam[3], iz, u, v, s[3], R[3][3], M[3][3], t0[3], t1[3], j0[3], j1[3];
    double dir[3], vlen, det;
} proj64;

typedef struct {
    uint8_t visible, uclamp, vclamp, detclamp;
} pflags;

struct sfo_bwd_ctx {
    sfo_view v;
    uint32_t n, chs, w, h, tw, tx, ty;
    uint32_t* ranges;
    uint32_t* gids;
    uint32_t ni;
    pflags* pf;
    uint8_t* dec;      /* per pixel, per tile-list position: 0 skip, 1 contribute, 2 contribute capped */
    size_t* dec_off;   /* per pixel offset into dec */
    proj64* pj;
};

typedef struct sfo_bwd_ctx sfo_bwd_ctx;

static double d_sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

#define C0 0.2820948
#define C1 0.48860252
#define C20 0.9461747
#define C21 0.31539157
#define C22 -1.0925485
#define C23 0.54627424
#define K9 -0.5900436

/* SH basis values of project.rs:342-388 in coefficient order (k = 0..15):
 * k -> planes (3k + c). */
static void sh_basis(double x, double y, double z, int chs, double B[16]) {
    memset(B, 0, 16 * sizeof(double));
    B[0] = C0;
    if (chs >= 4) { B[1] = -C1 * y; B[2] = C1 * z; B[3] = -C1 * x; }
    if (chs >= 9) {
        B[4] = C23 * 2.0 * x * y;       /* b7, planes 12.. */
        B[5] = C22 * z * y;             /* b5, planes 15.. */
        B[6] = C20 * z * z - C21;       /* b6, planes 18.. */
        B[7] = C22 * z * x;             /* b4, planes 21.. */
        B[8] = C23 * (x * x - y * y);   /* b8, planes 24.. */
    }
    if (chs >= 16) {
        double c1x = x * x - y * y, c1y = 2.0 * x * y;
        double t0 = -2.285229 * z * z + 0.4570458, t1 = 1.4453057 * z;
        B[9] = K9 * x * c1y + y * c1x;
        B[10] = K9 * x * c1x - y * c1y;
        B[11] = t0 * y;
        B[12] = t0 * x;
        B[13] = z * (1.8658817 * z * z - 1.119529);
        B[14] = t1 * c1y;
        B[15] = t1 * c1x;
    }
}

/* d basis_k / d(x, y, z). */
static void sh_basis_grad(double x, double y, double z, int chs, double D[16][3]) {
    memset(D, 0, 16 * 3 * sizeof(double));
    if (chs >= 4) { D[1][1] = -C1; D[2][2] = C1; D[3][0] = -C1; }
    if (chs >= 9) {
        D[4][0] = 2.0 * C23 * y; D[4][1] = 2.0 * C23 * x;
        D[5][1] = C22 * z; D[5][2] = C22 * y;
        D[6][2] = 2.0 * C20 * z;
        D[7][0] = C22 * z; D[7][2] = C22 * x;
        D[8][0] = 2.0 * C23 * x; D[8][1] = -2.0 * C23 * y;
    }
    if (chs >= 16) {
        double c1x = x * x - y * y, c1y = 2.0 * x * y;
        double t0 = -2.285229 * z * z + 0.4570458, t1 = 1.4453057 * z;
        D[9][0] = K9 * c1y + K9 * x * 2.0 * y + y * 2.0 * x;
        D[9][1] = K9 * x * 2.0 * x + c1x + y * (-2.0 * y);
        D[10][0] = K9 * c1x + K9 * x * 2.0 * x - y * 2.0 * y;
        D[10][1] = K9 * x * (-2.0 * y) - c1y - y * 2.0 * x;
        D[11][1] = t0; D[11][2] = -2.0 * 2.285229 * z * y;
        D[12][0] = t0; D[12][2] = -2.0 * 2.285229 * z * x;
        D[13][2] = 3.0 * 1.8658817 * z * z - 1.119529;
        D[14][0] = t1 * 2.0 * y; D[14][1] = t1 * 2.0 * x; D[14][2] = 1.4453057 * c1y;
        D[15][0] = t1 * 2.0 * x; D[15][1] = -t1 * 2.0 * y; D[15][2] = 1.4453057 * c1x;
    }
}

/* project_splats in f64 for splat i; returns visibility. */
static int proj_one(sfo_bwd_ctx* B, const double* A, const double* sh, uint32_t i, int record,
                    proj64* P) {
    const sfo_view* v = &B->v;
    size_t n = B->n;
    pflags* f = &B->pf[i];
    double mean[3] = {A[0 * n + i], A[1 * n + i], A[2 * n + i]};
    double qw = A[3 * n + i], qx = A[4 * n + i], qy = A[5 * n + i], qz = A[6 * n + i];
    for (int k = 0; k < 3; k++) P->s[k] = exp(A[(7 + k) * n + i]);
    P->op = d_sigmoid(A[10 * n + i]);
    double r0[3] = {v->rot0[0], v->rot0[1], v->rot0[2]};
    double r1[3] = {v->rot1[0], v->rot1[1], v->rot1[2]};
    double r2[3] = {v->rot2[0], v->rot2[1], v->rot2[2]};
    for (int k = 0; k < 3; k++) {
        const double* rr = k == 0 ? r0 : k == 1 ? r1 : r2;
        P->cam[k] = rr[0] * mean[0] + rr[1] * mean[1] + rr[2] * mean[2] + v->trans[k];
    }
    if (record) f->visible = !(P->op < 10.0 / 255.0) && !(P->cam[2] <= 0.1);
    if (!f->visible) return 0;
    P->iz = 1.0 / P->cam[2];
    double x2 = qx * qx, y2 = qy * qy, z2 = qz * qz, xy = qx * qy, xz = qx * qz, yz = qy * qz;
    double wx = qw * qx, wy = qw * qy, wz = qw * qz;
    double R[3][3] = {{1 - 2 * (y2 + z2), 2 * (xy - wz), 2 * (xz + wy)},
                      {2 * (xy + wz), 1 - 2 * (x2 + z2), 2 * (yz - wx)},
                      {2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (x2 + y2)}};
    memcpy(P->R, R, sizeof R);
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 3; k++) P->M[a][k] = R[a][k] * P->s[k];
    double limx = 1.3 * v->img[0] / (2.0 * v->focal[0]);
    double limy = 1.3 * v->img[1] / (2.0 * v->focal[1]);
    double u = P->cam[0] * P->iz, vv = P->cam[1] * P->iz;
    if (record) {
        f->uclamp = u < -limx || u > limx;
        f->vclamp = vv < -limy || vv > limy;
    }
    if (f->uclamp) u = u < 0 ? -limx : limx;
    if (f->vclamp) vv = vv < 0 ? -limy : limy;
    P->u = u;
    P->v = vv;
    double fxz = v->focal[0] * P->iz, fyz = v->focal[1] * P->iz;
    for (int k = 0; k < 3; k++) {
        P->t0[k] = fxz * r0[k] - fxz * u * r2[k];
        P->t1[k] = fyz * r1[k] - fyz * vv * r2[k];
    }
    for (int k = 0; k < 3; k++) {
        P->j0[k] = P->t0[0] * P->M[0][k] + P->t0[1] * P->M[1][k] + P->t0[2] * P->M[2][k];
        P->j1[k] = P->t1[0] * P->M[0][k] + P->t1[1] * P->M[1][k] + P->t1[2] * P->M[2][k];
    }
    P->ca = P->j0[0] * P->j0[0] + P->j0[1] * P->j0[1] + P->j0[2] * P->j0[2] + 0.3;
    P->cb = P->j0[0] * P->j1[0] + P->j0[1] * P->j1[1] + P->j0[2] * P->j1[2];
    P->cc = P->j1[0] * P->j1[0] + P->j1[1] * P->j1[1] + P->j1[2] * P->j1[2] + 0.3;
    double det = P->ca * P->cc - P->cb * P->cb;
    if (record) f->detclamp = det < 1e-6;
    if (f->detclamp) det = 1e-6;
    P->det = det;
    P->qa = P->cc / det;
    P->qb = -P->cb / det;
    P->qc = P->ca / det;
    double d[3] = {mean[0] - v->world_pos[0], mean[1] - v->world_pos[1], mean[2] - v->world_pos[2]};
    P->vlen = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; k++) P->dir[k] = d[k] / P->vlen;
    double Bs[16];
    sh_basis(P->dir[0], P->dir[1], P->dir[2], (int)B->chs, Bs);
    double col[3] = {0, 0, 0};
    for (uint32_t k = 0; k < B->chs; k++)
        for (int c = 0; c < 3; c++) col[c] += Bs[k] * sh[(size_t)(3 * k + c) * n + i];
    P->r = col[0] + 0.5;
    P->g = col[1] + 0.5;
    P->b = col[2] + 0.5;
    P->mx = v->focal[0] * P->cam[0] * P->iz + v->img[0] * 0.5;
    P->my = v->focal[1] * P->cam[1] * P->iz + v->img[1] * 0.5;
    return 1;
}

sfo_bwd_ctx* sfo_bwd_create(const sfo_view* v, uint32_t n, uint32_t sh_per_ch, uint32_t w,
                            uint32_t h, uint32_t tile_w, const uint32_t* ranges,
                            const uint32_t* gids, uint32_t num_isects) {
    sfo_bwd_ctx* B = (sfo_bwd_ctx*)calloc(1, sizeof *B);
    B->v = *v;
    B->n = n; B->chs = sh_per_ch; B->w = w; B->h = h; B->tw = tile_w;
    B->tx = (w + tile_w - 1) / tile_w;
    B->ty = (h + tile_w - 1) / tile_w;
    uint32_t T = B->tx * B->ty;
    B->ranges = (uint32_t*)malloc((size_t)2 * T * 4);
    memcpy(B->ranges, ranges, (size_t)2 * T * 4);
    B->ni = num_isects;
    B->gids = (uint32_t*)malloc((size_t)(num_isects ? num_isects : 1) * 4);
    memcpy(B->gids, gids, (size_t)num_isects * 4);
    B->pf = (pflags*)calloc(n, sizeof(pflags));
    B->pj = (proj64*)calloc(n, sizeof(proj64));
    B->dec_off = (size_t*)malloc((size_t)w * h * sizeof(size_t));
    size_t tot = 0;
    for (uint32_t py = 0; py < h; py++)
        for (uint32_t px = 0; px < w; px++) {
            uint32_t t = (px / tile_w) + (py / tile_w) * B->tx;
            B->dec_off[(size_t)py * w + px] = tot;
            tot += ranges[2 * t + 1] - ranges[2 * t];
        }
    B->dec = (uint8_t*)calloc(tot ? tot : 1, 1);
    return B;
}

void sfo_bwd_destroy(sfo_bwd_ctx* B) {
    if (!B) return;
    free(B->ranges); free(B->gids); free(B->pf); free(B->pj); free(B->dec_off); free(B->dec);
    free(B);
}

/* f64 forward + L2 loss.  out_rgba (optional) receives the image. */
double sfo_bwd_loss(sfo_bwd_ctx* B, const double* A, const double* sh, const float* target,
                    int record, double* out_rgba) {
    for (uint32_t i = 0; i < B->n; i++) proj_one(B, A, sh, i, record, &B->pj[i]);
    double loss = 0.0;
    for (uint32_t py = 0; py < B->h; py++)
        for (uint32_t px = 0; px < B->w; px++) {
            uint32_t t = (px / B->tw) + (py / B->tw) * B->tx;
            uint32_t s0 = B->ranges[2 * t], s1 = B->ranges[2 * t + 1];
            size_t p = (size_t)py * B->w + px;
            uint8_t* dec = B->dec + B->dec_off[p];
            double T = 1.0, c[3] = {0, 0, 0};
            double pxf = px + 0.5, pyf = py + 0.5;
            int done = 0;
            for (uint32_t k = s0; k < s1; k++) {
                uint8_t* d = &dec[k - s0];
                if (record) *d = 0;
                if (done) continue;
                proj64* P = &B->pj[B->gids[k]];
                double dx = P->mx - pxf, dy = P->my - pyf;
                double power = 0.5 * (P->qa * dx * dx + P->qc * dy * dy) + P->qb * dx * dy;
                double araw = P->op * exp(-power);
                if (record) {
                    if (araw >= 1.0 / 255.0) *d = araw > 0.999 ? 2 : 1;
                }
                if (!*d) continue;
                double alpha = *d == 2 ? 0.999 : araw;
                double vis = alpha * T;
                c[0] += P->r * vis; c[1] += P->g * vis; c[2] += P->b * vis;
                T *= 1.0 - alpha;
                if (record && T < 1.0 / 255.0) done = 1;
            }
            double img[4] = {c[0], c[1], c[2], 1.0 - T};
            for (int q = 0; q < 4; q++) {
                double e = img[q] - (double)target[4 * p + q];
                loss += e * e;
                if (out_rgba) out_rgba[4 * p + q] = img[q];
            }
        }
    return loss;
}

/* Projection adjoint of splat i (project.rs:234-388) given its screen-space
 * gradient G2 = d/d(mx, my, conic a, b, c, opacity, r, g, b); B->pj[i] and
 * B->pf[i] hold its f64 forward state. */
static void proj_adjoint(sfo_bwd_ctx* B, uint32_t i, const double* G2, const double* A,
                         const double* sh, double* gA, double* gS) {
    size_t n = B->n;
    const sfo_view* v = &B->v;
    if (!B->pf[i].visible) return;
    proj64* P = &B->pj[i];
    /* colour -> SH and view direction */
    double Bs[16], D[16][3];
    sh_basis(P->dir[0], P->dir[1], P->dir[2], (int)B->chs, Bs);
    sh_basis_grad(P->dir[0], P->dir[1], P->dir[2], (int)B->chs, D);
    double gdir[3] = {0, 0, 0};
    for (uint32_t k = 0; k < B->chs; k++)
        for (int c = 0; c < 3; c++) {
            double gc = G2[6 + c];
            gS[(size_t)(3 * k + c) * n + i] = Bs[k] * gc;
            double s = sh[(size_t)(3 * k + c) * n + i] * gc;
            gdir[0] += s * D[k][0]; gdir[1] += s * D[k][1]; gdir[2] += s * D[k][2];
        }
    double dd = gdir[0] * P->dir[0] + gdir[1] * P->dir[1] + gdir[2] * P->dir[2];
    double gmean[3];
    for (int k = 0; k < 3; k++) gmean[k] = (gdir[k] - P->dir[k] * dd) / P->vlen;
    /* opacity */
    gA[10 * n + i] = G2[5] * P->op * (1.0 - P->op);
    /* conic -> cov2d */
    double gx = G2[2], gy = G2[3], gz = G2[4];
    double a = P->ca, b = P->cb, c = P->cc, inv = 1.0 / P->det, inv2 = inv * inv;
    double ga, gb, gc;
    if (!B->pf[i].detclamp) {
        ga = gx * (-c * c * inv2) + gy * (b * c * inv2) + gz * (inv - a * c * inv2);
        gb = gx * (2.0 * b * c * inv2) + gy * (-inv - 2.0 * b * b * inv2) + gz * (2.0 * a * b * inv2);
        gc = gx * (inv - a * c * inv2) + gy * (a * b * inv2) + gz * (-a * a * inv2);
    } else {
        ga = gz * inv; gb = -gy * inv; gc = gx * inv;
    }
    double gj0[3], gj1[3];
    for (int k = 0; k < 3; k++) {
        gj0[k] = 2.0 * ga * P->j0[k] + gb * P->j1[k];
        gj1[k] = gb * P->j0[k] + 2.0 * gc * P->j1[k];
    }
    double gM[3][3], gt0[3] = {0, 0, 0}, gt1[3] = {0, 0, 0};
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) {
            gM[r][k] = P->t0[r] * gj0[k] + P->t1[r] * gj1[k];
            gt0[r] += P->M[r][k] * gj0[k];
            gt1[r] += P->M[r][k] * gj1[k];
        }
    double gR[3][3];
    for (int k = 0; k < 3; k++) {
        double gs = 0;
        for (int r = 0; r < 3; r++) {
            gs += P->R[r][k] * gM[r][k];
            gR[r][k] = gM[r][k] * P->s[k];
        }
        gA[(7 + k) * n + i] = gs * P->s[k];
    }
    double qw = A[3 * n + i], qx = A[4 * n + i], qy = A[5 * n + i], qz = A[6 * n + i];
    double gqw = 0, gqx = 0, gqy = 0, gqz = 0;
    gqy += gR[0][0] * (-4 * qy); gqz += gR[0][0] * (-4 * qz);
    gqx += gR[0][1] * 2 * qy; gqy += gR[0][1] * 2 * qx; gqw += gR[0][1] * -2 * qz; gqz += gR[0][1] * -2 * qw;
    gqx += gR[0][2] * 2 * qz; gqz += gR[0][2] * 2 * qx; gqw += gR[0][2] * 2 * qy; gqy += gR[0][2] * 2 * qw;
    gqx += gR[1][0] * 2 * qy; gqy += gR[1][0] * 2 * qx; gqw += gR[1][0] * 2 * qz; gqz += gR[1][0] * 2 * qw;
    gqx += gR[1][1] * (-4 * qx); gqz += gR[1][1] * (-4 * qz);
    gqy += gR[1][2] * 2 * qz; gqz += gR[1][2] * 2 * qy; gqw += gR[1][2] * -2 * qx; gqx += gR[1][2] * -2 * qw;
    gqx += gR[2][0] * 2 * qz; gqz += gR[2][0] * 2 * qx; gqw += gR[2][0] * -2 * qy; gqy += gR[2][0] * -2 * qw;
    gqy += gR[2][1] * 2 * qz; gqz += gR[2][1] * 2 * qy; gqw += gR[2][1] * 2 * qx; gqx += gR[2][1] * 2 * qw;
    gqx += gR[2][2] * (-4 * qx); gqy += gR[2][2] * (-4 * qy);
    gA[3 * n + i] = gqw; gA[4 * n + i] = gqx; gA[5 * n + i] = gqy; gA[6 * n + i] = gqz;
    /* t0/t1 -> camera space */
    double r0[3] = {v->rot0[0], v->rot0[1], v->rot0[2]};
    double r1[3] = {v->rot1[0], v->rot1[1], v->rot1[2]};
    double r2[3] = {v->rot2[0], v->rot2[1], v->rot2[2]};
    double fx = v->focal[0], fy = v->focal[1];
    double giz = 0, gu = 0, gv = 0;
    for (int k = 0; k < 3; k++) {
        giz += gt0[k] * fx * (r0[k] - P->u * r2[k]) + gt1[k] * fy * (r1[k] - P->v * r2[k]);
        gu += gt0[k] * (-fx * P->iz * r2[k]);
        gv += gt1[k] * (-fy * P->iz * r2[k]);
    }
    double gcam[3] = {0, 0, 0};
    if (!B->pf[i].uclamp) { gcam[0] += gu * P->iz; giz += gu * P->cam[0]; }
    if (!B->pf[i].vclamp) { gcam[1] += gv * P->iz; giz += gv * P->cam[1]; }
    /* mean2d */
    gcam[0] += G2[0] * fx * P->iz; giz += G2[0] * fx * P->cam[0];
    gcam[1] += G2[1] * fy * P->iz; giz += G2[1] * fy * P->cam[1];
    gcam[2] += giz * (-P->iz * P->iz);
    for (int k = 0; k < 3; k++)
        gA[(size_t)k * n + i] = gmean[k] + r0[k] * gcam[0] + r1[k] * gcam[1] + r2[k] * gcam[2];
}

/* Analytic gradient of sfo_bwd_loss at the recorded point (call after
 * sfo_bwd_loss(record = 1) with the same parameters). */
void sfo_bwd_grad(sfo_bwd_ctx* B, const double* A, const double* sh, const float* target,
                  double* gA, double* gS) {
    size_t n = B->n;
    memset(gA, 0, n * 11 * sizeof(double));
    memset(gS, 0, n * 3 * B->chs * sizeof(double));
    double* g2 = (double*)calloc(n * 9, sizeof(double)); /* mx my qa qb qc op r g b */
    double* img = (double*)malloc((size_t)B->w * B->h * 4 * sizeof(double));
    sfo_bwd_loss(B, A, sh, target, 0, img);
    for (uint32_t py = 0; py < B->h; py++)
        for (uint32_t px = 0; px < B->w; px++) {
            uint32_t t = (px / B->tw) + (py / B->tw) * B->tx;
            uint32_t s0 = B->ranges[2 * t], s1 = B->ranges[2 * t + 1];
            size_t p = (size_t)py * B->w + px;
            uint8_t* dec = B->dec + B->dec_off[p];
            double pxf = px + 0.5, pyf = py + 0.5;
            double dC[4];
            for (int q = 0; q < 4; q++) dC[q] = 2.0 * (img[4 * p + q] - (double)target[4 * p + q]);
            /* forward transmittances */
            uint32_t cnt = 0;
            for (uint32_t k = s0; k < s1; k++) cnt += dec[k - s0] ? 1 : 0;
            double Tfin = 1.0 - img[4 * p + 3];
            double T = Tfin, acc[3] = {0, 0, 0};
            for (uint32_t kk = s1; kk > s0; kk--) {
                uint32_t k = kk - 1;
                uint8_t d = dec[k - s0];
                if (!d) continue;
                uint32_t g = B->gids[k];
                proj64* P = &B->pj[g];
                double dx = P->mx - pxf, dy = P->my - pyf;
                double power = 0.5 * (P->qa * dx * dx + P->qc * dy * dy) + P->qb * dx * dy;
                double G = exp(-power);
                double alpha = d == 2 ? 0.999 : P->op * G;
                T = T / (1.0 - alpha);
                double w = alpha * T;
                double* G2 = &g2[(size_t)g * 9];
                G2[6] += w * dC[0]; G2[7] += w * dC[1]; G2[8] += w * dC[2];
                double da = T * ((P->r - acc[0]) * dC[0] + (P->g - acc[1]) * dC[1] +
                                 (P->b - acc[2]) * dC[2]) + Tfin / (1.0 - alpha) * dC[3];
                acc[0] = alpha * P->r + (1.0 - alpha) * acc[0];
                acc[1] = alpha * P->g + (1.0 - alpha) * acc[1];
                acc[2] = alpha * P->b + (1.0 - alpha) * acc[2];
                if (d == 1) {
                    double dpow = -da * alpha;
                    G2[5] += da * G;
                    G2[0] += dpow * (P->qa * dx + P->qb * dy);
                    G2[1] += dpow * (P->qc * dy + P->qb * dx);
                    G2[2] += dpow * 0.5 * dx * dx;
                    G2[3] += dpow * dx * dy;
                    G2[4] += dpow * 0.5 * dy * dy;
                }
            }
            (void)cnt;
        }
    for (uint32_t i = 0; i < n; i++) proj_adjoint(B, i, &g2[(size_t)i * 9], A, sh, gA, gS);
    free(g2);
    free(img);
}

/* ------------------------------------------------------------------------
 * CPU reference gradient of the f32 render: the adjoint evaluated from the
 * f32 forward state (projected rows, final T and the index past the last
 * contributor per pixel, as the compositor leaves them), per-pixel
 * transmittance recomputed back to front; then the projection adjoint in f64
 * per splat.  Multi-threaded over tile rows (raster) and splats (project);
 * per-thread partial sums are added in thread order, so the result does not
 * depend on scheduling. */

typedef struct {
    uint32_t n, w, h, tw, tx, ty0, ty1;
    const uint32_t *ranges, *gids, *n_contrib;
    const float *proj, *final_t, *d_rgba;
    float* g2; /* 9 planes x n, this thread's partial sums */
} rbwd_job;

static void* rbwd_worker(void* arg) {
    rbwd_job* J = (rbwd_job*)arg;
    size_t n = J->n;
    for (uint32_t ty = J->ty0; ty < J->ty1; ty++)
        for (uint32_t tx = 0; tx < J->tx; tx++) {
            uint32_t t = tx + ty * J->tx;
            uint32_t s0 = J->ranges[2 * t];
            for (uint32_t yy = 0; yy < J->tw; yy++)
                for (uint32_t xx = 0; xx < J->tw; xx++) {
                    uint32_t px = tx * J->tw + xx, py = ty * J->tw + yy;
                    if (px >= J->w || py >= J->h) continue;
                    size_t p = (size_t)py * J->w + px;
                    uint32_t last = J->n_contrib[p];
                    float Tfin = J->final_t[p], T = Tfin;
                    const float* dC = J->d_rgba + 4 * p;
                    float ar = 0.f, ag = 0.f, ab = 0.f;
                    float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
                    for (uint32_t kk = last; kk > s0; kk--) {
                        uint32_t k = kk - 1;
                        uint32_t g = J->gids[k];
                        const float* P = J->proj + (size_t)g * 9;
                        float dx = P[0] - pxf, dy = P[1] - pyf;
                        float power = 0.5f * (P[2] * dx * dx + P[4] * dy * dy) + P[3] * dx * dy;
                        float G = sfo_expf(-power);
                        float araw = P[8] * G;
                        float alpha = fminf(araw, 0.999f);
                        if (!(alpha >= 1.0f / 255.0f)) continue;
                        float om = 1.0f - alpha;
                        T = T / om;
                        float wgt = alpha * T;
                        float* o = J->g2;
                        o[6 * n + g] += wgt * dC[0];
                        o[7 * n + g] += wgt * dC[1];
                        o[8 * n + g] += wgt * dC[2];
                        float dal = T * ((P[5] - ar) * dC[0] + (P[6] - ag) * dC[1] +
                                         (P[7] - ab) * dC[2]) + Tfin / om * dC[3];
                        ar = alpha * P[5] + om * ar;
                        ag = alpha * P[6] + om * ag;
                        ab = alpha * P[7] + om * ab;
                        if (araw < 0.999f) {
                            float dpow = -dal * alpha;
                            o[5 * n + g] += dal * G;
                            o[0 * n + g] += dpow * (P[2] * dx + P[3] * dy);
                            o[1 * n + g] += dpow * (P[4] * dy + P[3] * dx);
                            o[2 * n + g] += dpow * 0.5f * dx * dx;
                            o[3 * n + g] += dpow * dx * dy;
                            o[4 * n + g] += dpow * 0.5f * dy * dy;
                        }
                    }
                }
        }
    return NULL;
}

typedef struct {
    const sfo_view* v;
    uint32_t n, chs, lo, hi, nthreads;
    const float *attrs, *sh;
    float** parts;
    float *gA, *gS;
} pbwd_job;

static void* pbwd_worker(void* arg) {
    pbwd_job* J = (pbwd_job*)arg;
    size_t n = J->n;
    sfo_bwd_ctx B;
    memset(&B, 0, sizeof B);
    B.v = *J->v;
    B.n = 1;
    B.chs = J->chs;
    pflags pf;
    proj64 pj;
    B.pf = &pf;
    B.pj = &pj;
    double A1[11], S1[48], gA1[11], gS1[48], G2[9];
    for (uint32_t i = J->lo; i < J->hi; i++) {
        for (int k = 0; k < 11; k++) A1[k] = J->attrs[(size_t)k * n + i];
        for (uint32_t k = 0; k < 3 * J->chs; k++) S1[k] = J->sh[(size_t)k * n + i];
        for (int k = 0; k < 9; k++) {
            double s = 0.0;
            for (uint32_t t = 0; t < J->nthreads; t++) s += J->parts[t][(size_t)k * n + i];
            G2[k] = s;
        }
        memset(&pf, 0, sizeof pf);
        memset(gA1, 0, sizeof gA1);
        memset(gS1, 0, sizeof gS1);
        if (proj_one(&B, A1, S1, 0, 1, &pj)) proj_adjoint(&B, 0, G2, A1, S1, gA1, gS1);
        for (int k = 0; k < 11; k++) J->gA[(size_t)k * n + i] = (float)gA1[k];
        for (uint32_t k = 0; k < 3 * J->chs; k++) J->gS[(size_t)k * n + i] = (float)gS1[k];
    }
    return NULL;
}

void sfo_backward(const sfo_view* v, const float* attrs, const float* sh, uint32_t n,
                  uint32_t sh_per_ch, uint32_t w, uint32_t h, uint32_t tile_w,
                  const uint32_t* ranges, const uint32_t* gids, const float* projected,
                  const float* final_t, const uint32_t* n_contrib, const float* d_rgba,
                  int nthreads, float* grad_attrs, float* grad_sh) {
    uint32_t tx = (w + tile_w - 1) / tile_w, ty = (h + tile_w - 1) / tile_w;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    if ((uint32_t)nthreads > ty) nthreads = (int)ty;
    pthread_t th[64];
    rbwd_job rj[64];
    float* parts[64];
    uint32_t per = (ty + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        parts[t] = (float*)calloc((size_t)n * 9, sizeof(float));
        rbwd_job* J = &rj[t];
        J->n = n; J->w = w; J->h = h; J->tw = tile_w; J->tx = tx;
        J->ty0 = (uint32_t)t * per < ty ? (uint32_t)t * per : ty;
        J->ty1 = J->ty0 + per < ty ? J->ty0 + per : ty;
        J->ranges = ranges; J->gids = gids; J->n_contrib = n_contrib; J->proj = projected;
        J->final_t = final_t; J->d_rgba = d_rgba; J->g2 = parts[t];
        pthread_create(&th[t], NULL, rbwd_worker, J);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pbwd_job pjob[64];
    uint32_t pper = (n + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        pbwd_job* J = &pjob[t];
        J->v = v; J->n = n; J->chs = sh_per_ch; J->nthreads = (uint32_t)nthreads;
        J->lo = (uint32_t)t * pper < n ? (uint32_t)t * pper : n;
        J->hi = J->lo + pper < n ? J->lo + pper : n;
        J->attrs = attrs; J->sh = sh; J->parts = parts; J->gA = grad_attrs; J->gS = grad_sh;
        pthread_create(&th[t], NULL, pbwd_worker, J);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    for (int t = 0; t < nthreads; t++) free(parts[t]);
}
