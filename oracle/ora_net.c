/*
 * ora_net.c -- CPU restatement of the learned Poisson initial-guess network
 * (TEST INFRASTRUCTURE ONLY; no reference counterpart).
 *
 * Same architecture, parameter layout and rounding points as the device
 * kernels:
 *   x0  = relu(conv(in16, 16->16))          in16 = (b, phi1, phi2, 0 x 13)
 *   x1  = relu(conv(pool(x0), 16->32))      x1b = relu(conv(x1, 32->32))
 *   x2  = relu(conv(pool(x1b), 32->64))     x2b = relu(conv(x2, 64->64))
 *   y1  = relu(conv([up(x2b), x1b], 96->32))
 *   y0  = relu(conv([up(y1), x0], 48->16))
 *   out = conv(y0, 16->1)
 * Parameters per layer: W[27][cin][cout] (tap t = 9*(dz+1) + 3*(dy+1) +
 * (dx+1)) followed by bias[cout].  Loss = mean (out - target)^2.
 *
 * round = 1 rounds weights, stored activations and stored activation
 * gradients to BF16 (round to nearest even) and the head output and its
 * gradient to FP32, where the device kernels do; arithmetic is FP64.
 * round = 0 is exact FP64 throughout (for the library and
 * finite-difference pins).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "tb_oracle.h"

#define NL 8
static const int CIN[NL] = {16, 16, 32, 32, 64, 96, 48, 16};
static const int COUT[NL] = {16, 32, 32, 64, 64, 32, 16, 1};

static double bf(double x) {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
        u = (u | 0x00400000u) & 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    memcpy(&f, &u, 4);
    return (double)f;
}

static void offsets(long *wo, long *bo) {
    long o = 0;
    for (int l = 0; l < NL; l++) {
        wo[l] = o;
        o += 27L * CIN[l] * COUT[l];
        bo[l] = o;
        o += COUT[l];
    }
}

int64_t ora_net_param_count(void) {
    long wo[NL], bo[NL];
    offsets(wo, bo);
    return bo[NL - 1] + COUT[NL - 1];
}

typedef struct {
    int nx, ny, nz;
} dims3;

static long vid(dims3 d, int x, int y, int z) { return ((long)z * d.ny + y) * d.nx + x; }

static int nb(dims3 d, int t, int x, int y, int z, int *xx, int *yy, int *zz) {
    *xx = x + t % 3 - 1;
    *yy = y + (t / 3) % 3 - 1;
    *zz = z + t / 9 - 1;
    return *xx >= 0 && *yy >= 0 && *zz >= 0 && *xx < d.nx && *yy < d.ny && *zz < d.nz;
}

/* value of input channel c of the (possibly fused) operand at voxel (x,y,z)
 * of the fine grid: channels [0, Ca) from A (half resolution when up),
 * [Ca, Ca+Cb) from B */
static double src_val(const double *A, int Ca, int up, const double *B, int Cb, dims3 d, int c, int x, int y, int z) {
    if (c < Ca) {
        if (up) {
            dims3 h = {d.nx / 2, d.ny / 2, d.nz / 2};
            return A[vid(h, x / 2, y / 2, z / 2) * Ca + c];
        }
        return A[vid(d, x, y, z) * Ca + c];
    }
    return B[vid(d, x, y, z) * Cb + (c - Ca)];
}

/* out[v][n] = act(bias + sum_t sum_c W[t][c][n] X[v+off(t)][c]) */
static void conv_fwd(const double *A, int Ca, int up, const double *B, int Cb, dims3 d, const double *W,
                     const double *bias, int cout, int relu, int round, double *out) {
    const int cin = Ca + Cb;
    long nv = (long)d.nx * d.ny * d.nz;
#pragma omp parallel for schedule(static)
    for (long v = 0; v < nv; v++) {
        int x = (int)(v % d.nx), y = (int)((v / d.nx) % d.ny), z = (int)(v / ((long)d.nx * d.ny));
        double acc[96];
        for (int n = 0; n < cout; n++) acc[n] = 0.0;
        for (int t = 0; t < 27; t++) {
            int xx, yy, zz;
            if (!nb(d, t, x, y, z, &xx, &yy, &zz)) continue;
            for (int c = 0; c < cin; c++) {
                double xv = src_val(A, Ca, up, B, Cb, d, c, xx, yy, zz);
                if (xv == 0.0) continue;
                const double *wr = W + ((long)t * cin + c) * cout;
                for (int n = 0; n < cout; n++) acc[n] += wr[n] * xv;
            }
        }
        for (int n = 0; n < cout; n++) {
            double s = acc[n] + (bias ? bias[n] : 0.0);
            if (relu && !(s > 0.0)) s = 0.0;
            out[v * cout + n] = round ? bf(s) : s;
        }
    }
}

/* dX[u][c] = sum_t sum_n W[t][c][n] dY[u - off(t)][n]  on the fused fine
 * grid (cin channels); optionally masked by m[u][c] > 0 */
static void conv_bwd_data(const double *dY, int cout, dims3 d, const double *W, int cin, const double *mask, int round,
                          double *dX) {
    long nv = (long)d.nx * d.ny * d.nz;
#pragma omp parallel for schedule(static)
    for (long u = 0; u < nv; u++) {
        int x = (int)(u % d.nx), y = (int)((u / d.nx) % d.ny), z = (int)(u / ((long)d.nx * d.ny));
        double acc[96];
        for (int c = 0; c < cin; c++) acc[c] = 0.0;
        for (int t = 0; t < 27; t++) {
            int xx, yy, zz;
            if (!nb(d, 26 - t, x, y, z, &xx, &yy, &zz)) continue;   /* v = u - off(t) */
            const double *g = dY + vid(d, xx, yy, zz) * cout;
            for (int c = 0; c < cin; c++) {
                const double *wr = W + ((long)t * cin + c) * cout;
                double s = 0.0;
                for (int n = 0; n < cout; n++) s += wr[n] * g[n];
                acc[c] += s;
            }
        }
        for (int c = 0; c < cin; c++) {
            double s = acc[c];
            if (mask && !(mask[u * cin + c] > 0.0)) s = 0.0;
            dX[u * cin + c] = round ? bf(s) : s;
        }
    }
}

/* dW[t][c][n] += sum_v X[v+off(t)][c] dY[v][n];  db[n] += sum_v dY[v][n] */
static void conv_wgrad(const double *A, int Ca, int up, const double *B, int Cb, dims3 d, const double *dY, int cout,
                       double *dW, double *db) {
    const int cin = Ca + Cb;
    long nv = (long)d.nx * d.ny * d.nz;
#pragma omp parallel for schedule(static)
    for (int tc = 0; tc < 27 * cin; tc++) {
        int t = tc / cin, c = tc % cin;
        double *acc = dW + (long)tc * cout;
        for (long v = 0; v < nv; v++) {
            int x = (int)(v % d.nx), y = (int)((v / d.nx) % d.ny), z = (int)(v / ((long)d.nx * d.ny));
            int xx, yy, zz;
            if (!nb(d, t, x, y, z, &xx, &yy, &zz)) continue;
            double xv = src_val(A, Ca, up, B, Cb, d, c, xx, yy, zz);
            if (xv == 0.0) continue;
            const double *g = dY + v * cout;
            for (int n = 0; n < cout; n++) acc[n] += xv * g[n];
        }
    }
    for (long v = 0; v < nv; v++)
        for (int n = 0; n < cout; n++) db[n] += dY[v * cout + n];
}

static void pool(const double *in, dims3 d, int C, int round, double *out) {
    dims3 h = {d.nx / 2, d.ny / 2, d.nz / 2};
    for (int Z = 0; Z < h.nz; Z++)
        for (int Y = 0; Y < h.ny; Y++)
            for (int X = 0; X < h.nx; X++)
                for (int c = 0; c < C; c++) {
                    double s = 0.0;
                    for (int q = 0; q < 8; q++)
                        s += in[vid(d, 2 * X + (q & 1), 2 * Y + ((q >> 1) & 1), 2 * Z + (q >> 2)) * C + c];
                    s *= 0.125;
                    out[vid(h, X, Y, Z) * C + c] = round ? bf(s) : s;
                }
}

/* gf = mask(gf + up(gc)/8) */
static void pool_bwd(const double *gc, double *gf, const double *mask, dims3 d, int C, int round) {
    dims3 h = {d.nx / 2, d.ny / 2, d.nz / 2};
    for (int z = 0; z < d.nz; z++)
        for (int y = 0; y < d.ny; y++)
            for (int x = 0; x < d.nx; x++)
                for (int c = 0; c < C; c++) {
                    long q = vid(d, x, y, z) * C + c;
                    double s = gf[q] + 0.125 * gc[vid(h, x / 2, y / 2, z / 2) * C + c];
                    if (!(mask[q] > 0.0)) s = 0.0;
                    gf[q] = round ? bf(s) : s;
                }
}

/* split the fused-input gradient du (Ca+Cb channels, fine grid d):
 * gA (half grid) = mask(sum of children), gB = du[., Ca:] */
static void up_bwd(const double *du, int Ca, int Cb, dims3 d, const double *maskA, int round, double *gA, double *gB) {
    const int C = Ca + Cb;
    dims3 h = {d.nx / 2, d.ny / 2, d.nz / 2};
    for (int Z = 0; Z < h.nz; Z++)
        for (int Y = 0; Y < h.ny; Y++)
            for (int X = 0; X < h.nx; X++)
                for (int c = 0; c < Ca; c++) {
                    double s = 0.0;
                    for (int q = 0; q < 8; q++)
                        s += du[vid(d, 2 * X + (q & 1), 2 * Y + ((q >> 1) & 1), 2 * Z + (q >> 2)) * C + c];
                    long o = vid(h, X, Y, Z) * Ca + c;
                    if (!(maskA[o] > 0.0)) s = 0.0;
                    gA[o] = round ? bf(s) : s;
                }
    long nv = (long)d.nx * d.ny * d.nz;
    for (long v = 0; v < nv; v++)
        for (int c = 0; c < Cb; c++) gB[v * Cb + c] = du[v * C + Ca + c];
}

typedef struct {
    dims3 d0, d1, d2;
    long v0, v1, v2;
    double *in16, *x0, *p0, *x1, *x1b, *p1, *x2, *x2b, *y1, *y0, *out;
    double *W[NL];   /* weights as used (rounded when round) */
    const double *bias[NL];
} net_fw;

static double *dz_alloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static int fw_run(net_fw *F, const double *w, const double *in3, int nx, int ny, int nz, int round) {
    if (nx % 4 || ny % 4 || nz % 4) return -1;
    long wo[NL], bo[NL];
    offsets(wo, bo);
    F->d0 = (dims3){nx, ny, nz};
    F->d1 = (dims3){nx / 2, ny / 2, nz / 2};
    F->d2 = (dims3){nx / 4, ny / 4, nz / 4};
    F->v0 = (long)nx * ny * nz;
    F->v1 = F->v0 / 8;
    F->v2 = F->v1 / 8;
    for (int l = 0; l < NL; l++) {
        long n = 27L * CIN[l] * COUT[l];
        F->W[l] = dz_alloc(n);
        for (long q = 0; q < n; q++) F->W[l][q] = round ? bf(w[wo[l] + q]) : w[wo[l] + q];
        F->bias[l] = w + bo[l];
    }
    F->in16 = dz_alloc(F->v0 * 16);
    for (long v = 0; v < F->v0; v++)
        for (int c = 0; c < 3; c++) F->in16[v * 16 + c] = round ? bf(in3[v * 3 + c]) : in3[v * 3 + c];
    F->x0 = dz_alloc(F->v0 * 16);
    F->p0 = dz_alloc(F->v1 * 16);
    F->x1 = dz_alloc(F->v1 * 32);
    F->x1b = dz_alloc(F->v1 * 32);
    F->p1 = dz_alloc(F->v2 * 32);
    F->x2 = dz_alloc(F->v2 * 64);
    F->x2b = dz_alloc(F->v2 * 64);
    F->y1 = dz_alloc(F->v1 * 32);
    F->y0 = dz_alloc(F->v0 * 16);
    F->out = dz_alloc(F->v0);
    conv_fwd(F->in16, 16, 0, NULL, 0, F->d0, F->W[0], F->bias[0], 16, 1, round, F->x0);
    pool(F->x0, F->d0, 16, round, F->p0);
    conv_fwd(F->p0, 16, 0, NULL, 0, F->d1, F->W[1], F->bias[1], 32, 1, round, F->x1);
    conv_fwd(F->x1, 32, 0, NULL, 0, F->d1, F->W[2], F->bias[2], 32, 1, round, F->x1b);
    pool(F->x1b, F->d1, 32, round, F->p1);
    conv_fwd(F->p1, 32, 0, NULL, 0, F->d2, F->W[3], F->bias[3], 64, 1, round, F->x2);
    conv_fwd(F->x2, 64, 0, NULL, 0, F->d2, F->W[4], F->bias[4], 64, 1, round, F->x2b);
    conv_fwd(F->x2b, 64, 1, F->x1b, 32, F->d1, F->W[5], F->bias[5], 32, 1, round, F->y1);
    conv_fwd(F->y1, 32, 1, F->x0, 16, F->d0, F->W[6], F->bias[6], 16, 1, round, F->y0);
    conv_fwd(F->y0, 16, 0, NULL, 0, F->d0, F->W[7], F->bias[7], 1, 0, 0, F->out);
    if (round)
        for (long v = 0; v < F->v0; v++) F->out[v] = (double)(float)F->out[v];
    return 0;
}

static void fw_free(net_fw *F) {
    for (int l = 0; l < NL; l++) free(F->W[l]);
    free(F->in16); free(F->x0); free(F->p0); free(F->x1); free(F->x1b); free(F->p1);
    free(F->x2); free(F->x2b); free(F->y1); free(F->y0); free(F->out);
}

int ora_net_forward(const double *w, const double *in3, int nx, int ny, int nz, int round, double *out) {
    net_fw F;
    memset(&F, 0, sizeof(F));
    if (fw_run(&F, w, in3, nx, ny, nz, round)) return -1;
    memcpy(out, F.out, sizeof(double) * F.v0);
    fw_free(&F);
    return 0;
}

int ora_net_gradient(const double *w, const double *in3, const double *target, int nx, int ny, int nz, int round,
                     double *loss, double *grad) {
    net_fw F;
    memset(&F, 0, sizeof(F));
    if (fw_run(&F, w, in3, nx, ny, nz, round)) return -1;
    long wo[NL], bo[NL];
    offsets(wo, bo);
    memset(grad, 0, sizeof(double) * (size_t)ora_net_param_count());
    const long v0 = F.v0, v1 = F.v1, v2 = F.v2;
    double *gout = dz_alloc(v0);
    double L = 0.0;
    for (long v = 0; v < v0; v++) {
        double dd = F.out[v] - target[v];
        L += dd * dd;
        gout[v] = 2.0 * dd / (double)v0;
        if (round) gout[v] = (double)(float)gout[v];
    }
    *loss = L / (double)v0;
    double *gy0 = dz_alloc(v0 * 16), *du0 = dz_alloc(v0 * 48), *gy1 = dz_alloc(v1 * 32), *gx0 = dz_alloc(v0 * 16);
    double *du1 = dz_alloc(v1 * 96), *gx2b = dz_alloc(v2 * 64), *gx1b = dz_alloc(v1 * 32), *gx2 = dz_alloc(v2 * 64);
    double *gp1 = dz_alloc(v2 * 32), *gx1 = dz_alloc(v1 * 32), *gp0 = dz_alloc(v1 * 16);
    /* head */
    conv_bwd_data(gout, 1, F.d0, F.W[7], 16, F.y0, round, gy0);
    conv_wgrad(F.y0, 16, 0, NULL, 0, F.d0, gout, 1, grad + wo[7], grad + bo[7]);
    /* y0 */
    conv_wgrad(F.y1, 32, 1, F.x0, 16, F.d0, gy0, 16, grad + wo[6], grad + bo[6]);
    conv_bwd_data(gy0, 16, F.d0, F.W[6], 48, NULL, round, du0);
    up_bwd(du0, 32, 16, F.d0, F.y1, round, gy1, gx0);
    /* y1 */
    conv_wgrad(F.x2b, 64, 1, F.x1b, 32, F.d1, gy1, 32, grad + wo[5], grad + bo[5]);
    conv_bwd_data(gy1, 32, F.d1, F.W[5], 96, NULL, round, du1);
    up_bwd(du1, 64, 32, F.d1, F.x2b, round, gx2b, gx1b);
    /* x2b, x2 */
    conv_wgrad(F.x2, 64, 0, NULL, 0, F.d2, gx2b, 64, grad + wo[4], grad + bo[4]);
    conv_bwd_data(gx2b, 64, F.d2, F.W[4], 64, F.x2, round, gx2);
    conv_wgrad(F.p1, 32, 0, NULL, 0, F.d2, gx2, 64, grad + wo[3], grad + bo[3]);
    conv_bwd_data(gx2, 64, F.d2, F.W[3], 32, NULL, round, gp1);
    pool_bwd(gp1, gx1b, F.x1b, F.d1, 32, round);
    /* x1b, x1 */
    conv_wgrad(F.x1, 32, 0, NULL, 0, F.d1, gx1b, 32, grad + wo[2], grad + bo[2]);
    conv_bwd_data(gx1b, 32, F.d1, F.W[2], 32, F.x1, round, gx1);
    conv_wgrad(F.p0, 16, 0, NULL, 0, F.d1, gx1, 32, grad + wo[1], grad + bo[1]);
    conv_bwd_data(gx1, 32, F.d1, F.W[1], 16, NULL, round, gp0);
    pool_bwd(gp0, gx0, F.x0, F.d0, 16, round);
    /* x0 */
    conv_wgrad(F.in16, 16, 0, NULL, 0, F.d0, gx0, 16, grad + wo[0], grad + bo[0]);
    free(gout); free(gy0); free(du0); free(gy1); free(gx0); free(du1); free(gx2b); free(gx1b);
    free(gx2); free(gp1); free(gx1); free(gp0);
    fw_free(&F);
    return 0;
}
