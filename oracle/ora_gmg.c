/*
 * ora_gmg.c -- FP64 geometric multigrid V-cycle on cell-centred levels, the
 * CPU counterpart of the device preconditioner (same hierarchy rule, same
 * red-black Gauss-Seidel smoother, same transfer operators).  Used by the
 * oracle's Poisson solve and as the CPU baseline's solver.
 * TEST INFRASTRUCTURE ONLY (see tb_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

#define COARSE_MAX 1024
#define COARSEST_CELLS 64   /* coarsening stops at this many cells */

static size_t lev_n(const ora_level *L) { return (size_t)L->nx * L->ny * L->nz; }

/* Coarsening rule: halve a direction when its spacing is at most half the
 * largest spacing (keeps coupling ratios <= 4 so point smoothing works);
 * if no direction qualifies, halve every coarsenable one. */
static int choose(const ora_level *L, int *cx, int *cy, int *cz) {
    int n[3] = {L->nx, L->ny, L->nz};
    double h[3] = {L->hx, L->hy, L->hz};
    double hmax = h[0];
    if (h[1] > hmax) hmax = h[1];
    if (h[2] > hmax) hmax = h[2];
    int c[3] = {0, 0, 0}, any = 0;
    for (int d = 0; d < 3; d++) {
        int can = (n[d] % 2 == 0) && (n[d] >= 4);
        if (can && 2.0 * h[d] <= hmax * (1.0 + 1e-9)) { c[d] = 1; any = 1; }
    }
    /* spacings within a factor 2 of each other: full coarsening */
    if (!any)
        for (int d = 0; d < 3; d++)
            if ((n[d] % 2 == 0) && (n[d] >= 4)) { c[d] = 1; any = 1; }
    *cx = c[0]; *cy = c[1]; *cz = c[2];
    return any;
}

void ora_gmg_free(ora_ctx *o) {
    for (int l = 0; l < o->nlev; l++) { free(o->lev[l].x); free(o->lev[l].b); free(o->lev[l].r); }
    free(o->coarse_inv);
    o->coarse_inv = NULL;
    o->nlev = 0;
}

static void apply_lev(const ora_level *L, const double *x, double *y);

int ora_gmg_build(ora_ctx *o) {
    ora_gmg_free(o);
    ora_level *L = &o->lev[0];
    L->nx = o->ncx; L->ny = o->ncy; L->nz = o->ncz;
    L->hx = o->cfg.lx / o->ncx; L->hy = o->cfg.ly / o->ncy; L->hz = o->cfg.lz / o->ncz;
    int nl = 1;
    while (nl < 16) {
        L = &o->lev[nl - 1];
        if (lev_n(L) <= COARSEST_CELLS) break;
        int cx, cy, cz;
        if (!choose(L, &cx, &cy, &cz)) break;
        L->cx = cx; L->cy = cy; L->cz = cz;
        ora_level *C = &o->lev[nl];
        C->nx = cx ? L->nx / 2 : L->nx; C->hx = cx ? 2.0 * L->hx : L->hx;
        C->ny = cy ? L->ny / 2 : L->ny; C->hy = cy ? 2.0 * L->hy : L->hy;
        C->nz = cz ? L->nz / 2 : L->nz; C->hz = cz ? 2.0 * L->hz : L->hz;
        nl++;
    }
    o->nlev = nl;
    for (int l = 0; l < nl; l++) {
        size_t n = lev_n(&o->lev[l]);
        o->lev[l].x = (double *)calloc(n, sizeof(double));
        o->lev[l].b = (double *)calloc(n, sizeof(double));
        o->lev[l].r = (double *)calloc(n, sizeof(double));
    }
    /* dense inverse of (A_c + J/n) on the coarsest level */
    L = &o->lev[nl - 1];
    int n = (int)lev_n(L);
    o->coarse_n = n;
    if (n <= COARSE_MAX) {
        double *A = (double *)calloc((size_t)n * n, sizeof(double));
        double *e = (double *)calloc((size_t)n, sizeof(double));
        double *col = (double *)calloc((size_t)n, sizeof(double));
        for (int a = 0; a < n; a++) {
            memset(e, 0, sizeof(double) * n);
            e[a] = 1.0;
            apply_lev(L, e, col);
            for (int r = 0; r < n; r++) A[(size_t)r * n + a] = col[r] + 1.0 / n;
        }
        /* Gauss-Jordan with partial pivoting into an identity */
        double *I = (double *)calloc((size_t)n * n, sizeof(double));
        for (int a = 0; a < n; a++) I[(size_t)a * n + a] = 1.0;
        for (int c = 0; c < n; c++) {
            int piv = c;
            for (int r = c + 1; r < n; r++) if (fabs(A[(size_t)r * n + c]) > fabs(A[(size_t)piv * n + c])) piv = r;
            if (piv != c)
                for (int k = 0; k < n; k++) {
                    double t = A[(size_t)c * n + k]; A[(size_t)c * n + k] = A[(size_t)piv * n + k]; A[(size_t)piv * n + k] = t;
                    t = I[(size_t)c * n + k]; I[(size_t)c * n + k] = I[(size_t)piv * n + k]; I[(size_t)piv * n + k] = t;
                }
            double d = 1.0 / A[(size_t)c * n + c];
            for (int k = 0; k < n; k++) { A[(size_t)c * n + k] *= d; I[(size_t)c * n + k] *= d; }
            for (int r = 0; r < n; r++) {
                if (r == c) continue;
                double f = A[(size_t)r * n + c];
                if (f == 0.0) continue;
                for (int k = 0; k < n; k++) { A[(size_t)r * n + k] -= f * A[(size_t)c * n + k]; I[(size_t)r * n + k] -= f * I[(size_t)c * n + k]; }
            }
        }
        o->coarse_inv = I;
        free(A); free(e); free(col);
    }
    return nl;
}

int ora_gmg_levels(ora_ctx *o, int32_t *dims, int cap) {
    if (o->nlev == 0) ora_gmg_build(o);
    for (int l = 0; l < o->nlev && l < cap; l++) {
        dims[3 * l] = o->lev[l].nx; dims[3 * l + 1] = o->lev[l].ny; dims[3 * l + 2] = o->lev[l].nz;
    }
    return o->nlev;
}

/* Neumann 7-point -Laplacian with constant spacing on a level */
static void apply_lev(const ora_level *L, const double *x, double *y) {
    int nx = L->nx, ny = L->ny, nz = L->nz;
    double cx = 1.0 / (L->hx * L->hx), cy = 1.0 / (L->hy * L->hy), cz = 1.0 / (L->hz * L->hz);
    size_t pl = (size_t)nx * ny;
#pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t id = ((size_t)k * ny + j) * nx + i;
                double xc = x[id], s = 0.0;
                if (i > 0) s += cx * (xc - x[id - 1]);
                if (i < nx - 1) s += cx * (xc - x[id + 1]);
                if (j > 0) s += cy * (xc - x[id - nx]);
                if (j < ny - 1) s += cy * (xc - x[id + nx]);
                if (k > 0) s += cz * (xc - x[id - pl]);
                if (k < nz - 1) s += cz * (xc - x[id + pl]);
                y[id] = s;
            }
}

static void rbgs(const ora_level *L, double *x, const double *b, int color) {
    int nx = L->nx, ny = L->ny, nz = L->nz;
    double cx = 1.0 / (L->hx * L->hx), cy = 1.0 / (L->hy * L->hy), cz = 1.0 / (L->hz * L->hz);
    size_t pl = (size_t)nx * ny;
#pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = (j + k + color) & 1; i < nx; i += 2) {
                size_t id = ((size_t)k * ny + j) * nx + i;
                double s = b[id], dg = 0.0;
                if (i > 0) { s += cx * x[id - 1]; dg += cx; }
                if (i < nx - 1) { s += cx * x[id + 1]; dg += cx; }
                if (j > 0) { s += cy * x[id - nx]; dg += cy; }
                if (j < ny - 1) { s += cy * x[id + nx]; dg += cy; }
                if (k > 0) { s += cz * x[id - pl]; dg += cz; }
                if (k < nz - 1) { s += cz * x[id + pl]; dg += cz; }
                x[id] = s / dg;
            }
}

static const double RW[4] = {0.125, 0.375, 0.375, 0.125};

static inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

/* restriction: adjoint of the prolongation divided by 2 per halved axis */
static void restrict_lev(const ora_level *F, const ora_level *C, const double *rf, double *bc) {
    int nxf = F->nx, nyf = F->ny, nzf = F->nz;
#pragma omp parallel for schedule(static) collapse(2)
    for (int K = 0; K < C->nz; K++)
        for (int J = 0; J < C->ny; J++)
            for (int I = 0; I < C->nx; I++) {
                double s = 0.0;
                int tk0 = F->cz ? -1 : 0, tk1 = F->cz ? 2 : 0;
                int tj0 = F->cy ? -1 : 0, tj1 = F->cy ? 2 : 0;
                int ti0 = F->cx ? -1 : 0, ti1 = F->cx ? 2 : 0;
                for (int tk = tk0; tk <= tk1; tk++) {
                    int kf = F->cz ? clampi(2 * K + tk, 0, nzf - 1) : K;
                    double wk = F->cz ? RW[tk + 1] : 1.0;
                    for (int tj = tj0; tj <= tj1; tj++) {
                        int jf = F->cy ? clampi(2 * J + tj, 0, nyf - 1) : J;
                        double wj = F->cy ? RW[tj + 1] : 1.0;
                        for (int ti = ti0; ti <= ti1; ti++) {
                            int ifn = F->cx ? clampi(2 * I + ti, 0, nxf - 1) : I;
                            double wi = F->cx ? RW[ti + 1] : 1.0;
                            s += wk * wj * wi * rf[((size_t)kf * nyf + jf) * nxf + ifn];
                        }
                    }
                }
                bc[((size_t)K * C->ny + J) * C->nx + I] = s;
            }
}

/* prolongation: piecewise-linear (3/4, 1/4) per halved axis, reflecting */
static void prolong_add(const ora_level *F, const ora_level *C, const double *xc, double *xf) {
    int nxc = C->nx, nyc = C->ny, nzc = C->nz;
#pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < F->nz; k++)
        for (int j = 0; j < F->ny; j++)
            for (int i = 0; i < F->nx; i++) {
                int K0, K1, J0, J1, I0, I1;
                double wk0, wk1, wj0, wj1, wi0, wi1;
                if (F->cz) { K0 = k >> 1; K1 = clampi((k & 1) ? K0 + 1 : K0 - 1, 0, nzc - 1); wk0 = 0.75; wk1 = 0.25; }
                else { K0 = K1 = k; wk0 = 1.0; wk1 = 0.0; }
                if (F->cy) { J0 = j >> 1; J1 = clampi((j & 1) ? J0 + 1 : J0 - 1, 0, nyc - 1); wj0 = 0.75; wj1 = 0.25; }
                else { J0 = J1 = j; wj0 = 1.0; wj1 = 0.0; }
                if (F->cx) { I0 = i >> 1; I1 = clampi((i & 1) ? I0 + 1 : I0 - 1, 0, nxc - 1); wi0 = 0.75; wi1 = 0.25; }
                else { I0 = I1 = i; wi0 = 1.0; wi1 = 0.0; }
#define XC(K, J, I) xc[((size_t)(K) * nyc + (J)) * nxc + (I)]
                double s = wk0 * (wj0 * (wi0 * XC(K0, J0, I0) + wi1 * XC(K0, J0, I1)) +
                                  wj1 * (wi0 * XC(K0, J1, I0) + wi1 * XC(K0, J1, I1))) +
                           wk1 * (wj0 * (wi0 * XC(K1, J0, I0) + wi1 * XC(K1, J0, I1)) +
                                  wj1 * (wi0 * XC(K1, J1, I0) + wi1 * XC(K1, J1, I1)));
#undef XC
                xf[((size_t)k * F->ny + j) * F->nx + i] += s;
            }
}

static void remove_mean(double *x, size_t n) {
    double m = 0.0;
    for (size_t a = 0; a < n; a++) m += x[a];
    m /= (double)n;
    for (size_t a = 0; a < n; a++) x[a] -= m;
}

static void vcycle_level(ora_ctx *o, int l) {
    ora_level *L = &o->lev[l];
    size_t n = lev_n(L);
    memset(L->x, 0, sizeof(double) * n);
    if (l == o->nlev - 1) {
        if (o->coarse_inv) {
            double *bb = (double *)malloc(sizeof(double) * n);
            memcpy(bb, L->b, sizeof(double) * n);
            remove_mean(bb, n);
            for (size_t r = 0; r < n; r++) {
                double s = 0.0;
                for (size_t c = 0; c < n; c++) s += o->coarse_inv[r * n + c] * bb[c];
                L->x[r] = s;
            }
            remove_mean(L->x, n);
            free(bb);
        } else {
            for (int s = 0; s < 50; s++) { rbgs(L, L->x, L->b, 0); rbgs(L, L->x, L->b, 1); }
            for (int s = 0; s < 50; s++) { rbgs(L, L->x, L->b, 1); rbgs(L, L->x, L->b, 0); }
        }
        return;
    }
    int pre = o->cfg.mg_pre > 0 ? o->cfg.mg_pre : 2, post = o->cfg.mg_post > 0 ? o->cfg.mg_post : 2;
    for (int s = 0; s < pre; s++) { rbgs(L, L->x, L->b, 0); rbgs(L, L->x, L->b, 1); }
    apply_lev(L, L->x, L->r);
    for (size_t a = 0; a < n; a++) L->r[a] = L->b[a] - L->r[a];
    restrict_lev(L, &o->lev[l + 1], L->r, o->lev[l + 1].b);
    vcycle_level(o, l + 1);
    prolong_add(L, &o->lev[l + 1], o->lev[l + 1].x, L->x);
    for (int s = 0; s < post; s++) { rbgs(L, L->x, L->b, 1); rbgs(L, L->x, L->b, 0); }
}

/* exported V-cycle on compact (nx*ny*nz, x fastest) cell arrays */
int ora_gmg_vcycle(ora_ctx *o, const double *r, double *z) {
    ora_vcycle(o, r, z);
    return 0;
}

/* z = V(r): one symmetric V-cycle, output projected to zero mean */
void ora_vcycle(ora_ctx *o, const double *r, double *z) {
    if (o->nlev == 0) ora_gmg_build(o);
    ora_level *L = &o->lev[0];
    size_t n = lev_n(L);
    memcpy(L->b, r, sizeof(double) * n);
    vcycle_level(o, 0);
    memcpy(z, L->x, sizeof(double) * n);
    remove_mean(z, n);
}
