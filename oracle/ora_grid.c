/*
 * ora_grid.c -- oracle restatement of the MAC grid, the signed-distance
 * reinitialisation, node classification and immersed-boundary setup.
 * TEST INFRASTRUCTURE ONLY (see tb_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

static double *dalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* Grid_generate_mesh_coordinates (Grid.c:470-591): coordinates accumulated by
 * repeated addition; Grid_compute_metric_coefficients (Grid.c:595-708). */
int ora_grid_init(ora_ctx *o) {
    int NX = o->NX, NY = o->NY, NZ = o->NZ;
    o->xc = dalloc(NX); o->yc = dalloc(NY); o->zc = dalloc(NZ);
    o->xu = dalloc(NX); o->yu = dalloc(NY); o->zu = dalloc(NZ);
    o->xv = dalloc(NX); o->yv = dalloc(NY); o->zv = dalloc(NZ);
    o->xw = dalloc(NX); o->yw = dalloc(NY); o->zw = dalloc(NZ);
    o->dx_c = dalloc(NX); o->dy_c = dalloc(NY); o->dz_c = dalloc(NZ);
    o->mxc = dalloc(NX); o->myc = dalloc(NY); o->mzc = dalloc(NZ);
    o->mxu = dalloc(NX); o->myv = dalloc(NY); o->mzw = dalloc(NZ);

    o->dEta = 1.0 / (double)o->ncx;
    o->dXi = 1.0 / (double)o->ncy;
    o->dPhi = 1.0 / (double)o->ncz;
    o->dx = o->cfg.lx * o->dEta;
    o->dy = o->cfg.ly * o->dXi;
    o->dz = o->cfg.lz * o->dPhi;
    double dx = o->dx, dy = o->dy, dz = o->dz;
    const double xmin = 0.0, ymin = 0.0, zmin = 0.0;
    double g;
    int i, j, k;
    /* 'u' */
    g = xmin; for (i = 0; i < NX; i++) { o->xu[i] = g; g += dx; }
    g = ymin + 0.5 * dy; for (j = 0; j < NY; j++) { o->yu[j] = g; g += dy; }
    g = zmin + 0.5 * dz; for (k = 0; k < NZ; k++) { o->zu[k] = g; g += dz; }
    /* 'v' */
    g = xmin + 0.5 * dx; for (i = 0; i < NX; i++) { o->xv[i] = g; g += dx; }
    g = ymin; for (j = 0; j < NY; j++) { o->yv[j] = g; g += dy; }
    g = zmin + 0.5 * dz; for (k = 0; k < NZ; k++) { o->zv[k] = g; g += dz; }
    /* 'w' */
    g = xmin + 0.5 * dx; for (i = 0; i < NX; i++) { o->xw[i] = g; g += dx; }
    g = ymin + 0.5 * dy; for (j = 0; j < NY; j++) { o->yw[j] = g; g += dy; }
    g = zmin; for (k = 0; k < NZ; k++) { o->zw[k] = g; g += dz; }
    /* 'c' */
    g = xmin + 0.5 * dx; for (i = 0; i < NX; i++) { o->xc[i] = g; g += dx; }
    g = ymin + 0.5 * dy; for (j = 0; j < NY; j++) { o->yc[j] = g; g += dy; }
    g = zmin + 0.5 * dz; for (k = 0; k < NZ; k++) { o->zc[k] = g; g += dz; }

    double dEta = o->dEta, dXi = o->dXi, dPhi = o->dPhi;
    for (i = 1; i < NX - 1; i++) { o->mxu[i] = dEta / (o->xc[i] - o->xc[i - 1]); o->dx_c[i] = o->xu[i + 1] - o->xu[i]; }
    o->dx_c[0] = o->xu[1] - o->xu[0];
    o->dx_c[NX - 1] = o->xu[NX - 1] - o->xu[NX - 2];
    o->mxu[0] = 0.5 * dEta / (o->xc[0] - o->xu[0]);
    o->mxu[NX - 1] = 0.5 * dEta / (o->xu[NX - 1] - o->xc[NX - 2]);

    for (j = 1; j < NY - 1; j++) { o->myv[j] = dXi / (o->yc[j] - o->yc[j - 1]); o->dy_c[j] = o->yv[j + 1] - o->yv[j]; }
    o->dy_c[0] = o->yv[1] - o->yv[0];
    o->dy_c[NY - 1] = o->yv[NY - 1] - o->yv[NY - 2];
    o->myv[0] = 0.5 * dXi / (o->yc[0] - o->yv[0]);
    o->myv[NY - 1] = 0.5 * dXi / (o->yv[NY - 1] - o->yc[NY - 2]);

    for (k = 1; k < NZ - 1; k++) { o->mzw[k] = dPhi / (o->zc[k] - o->zc[k - 1]); o->dz_c[k] = o->zw[k + 1] - o->zw[k]; }
    o->dz_c[0] = o->zw[1] - o->zw[0];
    o->dz_c[NZ - 1] = o->zw[NZ - 1] - o->zw[NZ - 2];
    o->mzw[0] = 0.5 * dPhi / (o->zc[0] - o->zw[0]);
    o->mzw[NZ - 1] = 0.5 * dPhi / (o->zw[NZ - 1] - o->zc[NZ - 2]);

    for (i = 0; i < NX - 1; i++) o->mxc[i] = dEta / (o->xu[i + 1] - o->xu[i]);
    o->mxc[NX - 1] = dEta / (o->xc[NX - 1] - o->xc[NX - 2]);
    for (j = 0; j < NY - 1; j++) o->myc[j] = dXi / (o->yv[j + 1] - o->yv[j]);
    o->myc[NY - 1] = dXi / (o->yc[NY - 1] - o->yc[NY - 2]);
    /* Grid.c:697 uses zc (not zw) for the z metric at cell centres */
    for (k = 0; k < NZ - 1; k++) o->mzc[k] = dPhi / (o->zc[k + 1] - o->zc[k]);
    o->mzc[NZ - 1] = dPhi / (o->zc[NZ - 1] - o->zc[NZ - 2]);
    return 0;
}

int ora_get_grid_array(ora_ctx *o, int which, double *out) {
    double *a[21] = {o->xc, o->yc, o->zc, o->xu, o->yu, o->zu, o->xv, o->yv, o->zv,
                     o->xw, o->yw, o->zw, o->dx_c, o->dy_c, o->dz_c,
                     o->mxc, o->myc, o->mzc, o->mxu, o->myv, o->mzw};
    int len[21] = {o->NX, o->NY, o->NZ, o->NX, o->NY, o->NZ, o->NX, o->NY, o->NZ,
                   o->NX, o->NY, o->NZ, o->NX, o->NY, o->NZ,
                   o->NX, o->NY, o->NZ, o->NX, o->NY, o->NZ};
    if (which < 0 || which > 20) return -1;
    memcpy(out, a[which], sizeof(double) * (size_t)len[which]);
    return len[which];
}

void ora_q_coords(const ora_ctx *o, int q, const double **xq, const double **yq,
                  const double **zq) {
    switch (q) {
    case TB_Q_U: *xq = o->xu; *yq = o->yu; *zq = o->zu; break;
    case TB_Q_V: *xq = o->xv; *yq = o->yv; *zq = o->zv; break;
    case TB_Q_W: *xq = o->xw; *yq = o->yw; *zq = o->zw; break;
    default: *xq = o->xc; *yq = o->yc; *zq = o->zc; break;
    }
}

/* Grid_get_*_status (Grid.c:1126-1224): OUTSIDE beyond [0, NX-1] etc. */
int ora_status(const ora_ctx *o, int q, int i, int j, int k) {
    if (i < 0 || i > o->NX - 1 || j < 0 || j > o->NY - 1 || k < 0 || k > o->NZ - 1)
        return TB_OUTSIDE;
    return (int)o->st[q][ora_idx(o, i, j, k)];
}

/* MyMath_interpolate_quantity, N == 2 branch (MyMath.c:24-41) */
double ora_interp2(double q0, double x0, double q1, double x1, double xi) {
    double m = (q1 - q0) / (x1 - x0);
    if (fabs(x1 - x0) < 1.0e-8) return q0;
    return m * (xi - x0) + q0;
}

/* MyMath_interpolate_quantity, N == 4 branch (MyMath.c:52-73).  The weight
 * measured from corner 0 multiplies corner 0 (exact only at mid points, as
 * in the reference). */
double ora_interp4(const double q[4], const double X1[4], const double X2[4],
                   double x1, double x2) {
    double a_Dx = 1.0 / (X1[1] - X1[0]);
    double Wp0 = fabs((x1 - X1[0]) * a_Dx);
    double Wp1 = 1.0 - Wp0;
    double q_top = Wp0 * q[2] + Wp1 * q[3];
    double q_bottom = Wp0 * q[0] + Wp1 * q[1];
    double a_Dy = 1.0 / (X2[2] - X2[0]);
    Wp0 = fabs((x2 - X2[0]) * a_Dy);
    double Wp2 = 1.0 - Wp0;
    return Wp0 * q_bottom + Wp2 * q_top;
}

/* ---------------------------------------------------------------------- */
/* Signed distance: Surface_initialize (Surface.c:96-192) + Levelset_*     */

static int sgn(double a) { return (a > 0.0) ? 1 : ((a < 0.0) ? -1 : 0); }

static int levelset_reinit(ora_ctx *o, const double *p0, double *p, int max_iterations) {
    int NX = o->NX, NY = o->NY, NZ = o->NZ;
    int NI = NX - 1, NJ = NY - 1, NK = NZ - 1;
    size_t NT = o->NT;
    double *pold = dalloc(NT);   /* VecDuplicate: zero initialised */
    double *Dx = dalloc(NT), *Dy = dalloc(NT), *Dz = dalloc(NT);
    double a_dEta = 1.0 / o->dEta, a_dXi = 1.0 / o->dXi, a_dPhi = 1.0 / o->dPhi;
    memcpy(p, p0, sizeof(double) * NT);

    /* Levelset_compute_fictitious_time (Levelset.c:590-637) */
    double dx_min = 1000.0, dy_min = 1000.0, dz_min = 1000.0;
    for (int i = 0; i < NI; i++) if (o->dx_c[i] < dx_min) dx_min = o->dx_c[i];
    for (int j = 0; j < NJ; j++) if (o->dy_c[j] < dy_min) dy_min = o->dy_c[j];
    for (int k = 0; k < NK; k++) if (o->dz_c[k] < dz_min) dz_min = o->dz_c[k];
    double d_min = ORA_MIN(dx_min, dy_min);
    d_min = ORA_MIN(d_min, dz_min);
    double dt = 0.5 * d_min;

    int converged = 0, c = 0;
    while (!converged) {
        /* Levelset_compute_derivatives (Levelset.c:163-261) */
#pragma omp parallel for schedule(static)
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++) {
                    size_t id = ora_idx(o, i, j, k);
                    if (i < NI - 1) Dx[id] = (p[id + 1] - p[id]) * o->mxu[i] * a_dEta;
                    else Dx[id] = (p[id] - p[id - 1]) * o->mxu[i - 1] * a_dEta;
                    if (j < NJ - 1) Dy[id] = (p[id + NX] - p[id]) * o->myv[j] * a_dXi;
                    else Dy[id] = (p[id] - p[id - NX]) * o->myv[j - 1] * a_dXi;
                    size_t pl = (size_t)NX * NY;
                    if (k < NK - 1) Dz[id] = (p[id + pl] - p[id]) * o->mzw[k] * a_dPhi;
                    else Dz[id] = (p[id] - p[id - pl]) * o->mzw[k - 1] * a_dPhi;
                }
        /* Levelset_do_Godunov (Levelset.c:267-494) */
#pragma omp parallel for schedule(static)
        for (int k = 0; k < NK; k++)
            for (int j = 0; j < NJ; j++)
                for (int i = 0; i < NI; i++) {
                    size_t id = ora_idx(o, i, j, k);
                    double dx = o->dx_c[i];
                    double sign_0 = p0[id] / sqrt(p0[id] * p0[id] + dx * dx);
                    double Dxp = Dx[id], Dxm = Dx[ora_idx(o, ORA_MAX(0, i - 1), j, k)];
                    double Dyp = Dy[id], Dym = Dy[ora_idx(o, i, ORA_MAX(0, j - 1), k)];
                    double Dzp = Dz[id], Dzm = Dz[ora_idx(o, i, j, ORA_MAX(0, k - 1))];
                    double ap = ORA_MAX(Dxm, 0.0), am = ORA_MIN(Dxm, 0.0);
                    double bp = ORA_MAX(Dxp, 0.0), bm = ORA_MIN(Dxp, 0.0);
                    double cp = ORA_MAX(Dym, 0.0), cm = ORA_MIN(Dym, 0.0);
                    double dp = ORA_MAX(Dyp, 0.0), dm = ORA_MIN(Dyp, 0.0);
                    double ep = ORA_MAX(Dzm, 0.0), em = ORA_MIN(Dzm, 0.0);
                    double fp = ORA_MAX(Dzp, 0.0), fm = ORA_MIN(Dzp, 0.0);
                    double tx, ty, tz;
                    if (sign_0 > 0.0) {
                        tx = ORA_MAX(ap * ap, bm * bm);
                        ty = ORA_MAX(cp * cp, dm * dm);
                        tz = ORA_MAX(ep * ep, fm * fm);
                    } else {
                        tx = ORA_MAX(am * am, bp * bp);
                        ty = ORA_MAX(cm * cm, dp * dp);
                        tz = ORA_MAX(em * em, fp * fp);
                    }
                    double H = sign_0 * (sqrt(tx + ty + tz) - 1.0);
                    double p_p = p0[id];
                    int ix;
                    ix = ORA_MIN(i + 1, NI - 1);
                    double p_e = p0[ora_idx(o, ix, j, k)];
                    double DeltaX = o->xc[ix];
                    ix = ORA_MIN(j + 1, NJ - 1);
                    double p_n = p0[ora_idx(o, i, ix, k)];
                    double DeltaY = o->yc[ix];
                    ix = ORA_MIN(k + 1, NK - 1);
                    double p_f = p0[ora_idx(o, i, j, ix)];
                    double DeltaZ = o->zc[ix];
                    ix = ORA_MAX(i - 1, 0);
                    double p_w = p0[ora_idx(o, ix, j, k)];
                    DeltaX -= o->xc[ix];
                    ix = ORA_MAX(j - 1, 0);
                    double p_s = p0[ora_idx(o, i, ix, k)];
                    DeltaY -= o->yc[ix];
                    ix = ORA_MAX(k - 1, 0);
                    double p_b = p0[ora_idx(o, i, j, ix)];
                    DeltaZ -= o->zc[ix];
                    if (((p_p * p_w) < 0.0) || ((p_p * p_e) < 0.0) || ((p_p * p_n) < 0.0) ||
                        ((p_p * p_s) < 0.0) || ((p_p * p_b) < 0.0) || ((p_p * p_f) < 0.0)) {
                        const double tol = 0.01;
                        /* TB_REINIT_SCALED: the stencil-width guard is
                         * min(tol, 1.5 d_c), so an interior central difference
                         * (width 2 d_c) is never dropped; identical to the
                         * reference's absolute tol whenever every d_c >= tol/1.5 */
                        double gx = tol, gy = tol, gz = tol;
                        if (o->cfg.reinit_guard == TB_REINIT_SCALED) {
                            gx = ORA_MIN(tol, 1.5 * o->dx_c[i]);
                            gy = ORA_MIN(tol, 1.5 * o->dy_c[j]);
                            gz = ORA_MIN(tol, 1.5 * o->dz_c[k]);
                        }
                        double Dphi_dx = (fabs(DeltaX) > gx) ? (p_e - p_w) / DeltaX : 0.0;
                        double Dphi_dy = (fabs(DeltaY) > gy) ? (p_n - p_s) / DeltaY : 0.0;
                        double Dphi_dz = (fabs(DeltaZ) > gz) ? (p_f - p_b) / DeltaZ : 0.0;
                        double Dt = sqrt(Dphi_dx * Dphi_dx + Dphi_dy * Dphi_dy + Dphi_dz * Dphi_dz);
                        if (Dt < tol) Dt = tol;
                        double D = p_p / Dt;
                        if (fabs(sign_0) < 1.0e-8) sign_0 = 0.0;
                        else if (sign_0 > 0.0) sign_0 = 1.0;
                        else sign_0 = -1.0;
                        double ddx = o->dx_c[i], ddy = o->dy_c[j], ddz = o->dz_c[k];
                        double Ds = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                        H = (sign_0 * fabs(p[id]) - D) / Ds;
                    }
                    p[id] += -H * dt;
                }
        /* Levelset_is_converged, NORM2 (Levelset.c:499-576): sequential sum */
        double E2 = 0.0;
        for (int k = 0; k < NZ; k++) {
            double dz = o->dz_c[k];
            for (int j = 0; j < NY; j++) {
                double dy = o->dy_c[j];
                for (int i = 0; i < NX; i++) {
                    double dx = o->dx_c[i];
                    size_t id = ora_idx(o, i, j, k);
                    double error = fabs(p[id] - pold[id]);
                    E2 += error * error * dx * dy * dz;
                }
            }
        }
        double W_E = sqrt(E2);
        int flag = (W_E < 1.0e-5);
        if ((c > max_iterations + 1) || flag) converged = 1;
        memcpy(pold, p, sizeof(double) * NT);
        c++;
    }
    /* Levelset_update_last_planes (Levelset.c:642-709) */
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++) p[ora_idx(o, NX - 1, j, k)] = p[ora_idx(o, NX - 2, j, k)];
    for (int k = 0; k < NZ; k++)
        for (int i = 0; i < NX; i++) p[ora_idx(o, i, NY - 1, k)] = p[ora_idx(o, i, NY - 2, k)];
    for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) p[ora_idx(o, i, j, NZ - 1)] = p[ora_idx(o, i, j, NZ - 2)];
    free(pold); free(Dx); free(Dy); free(Dz);
    return c;
}

/* Surface_find_q_sdf (Surface.c:216-334) */
static void find_q_sdf(ora_ctx *o) {
    double *c_sdf = o->f[TB_F_SDF_C];
    double *u_sdf = o->f[TB_F_SDF_U], *v_sdf = o->f[TB_F_SDF_V], *w_sdf = o->f[TB_F_SDF_W];
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                double q1, x1;
                if (i > 0) { q1 = c_sdf[ora_idx(o, i - 1, j, k)]; x1 = o->xc[i - 1]; }
                else { q1 = c_sdf[id]; x1 = o->xc[i]; }
                u_sdf[id] = ora_interp2(c_sdf[id], o->xc[i], q1, x1, o->xu[i]);
                if (j > 0) { q1 = c_sdf[ora_idx(o, i, j - 1, k)]; x1 = o->yc[j - 1]; }
                else { q1 = c_sdf[id]; x1 = o->yc[j]; }
                v_sdf[id] = ora_interp2(c_sdf[id], o->yc[j], q1, x1, o->yv[j]);
                if (k > 0) { q1 = c_sdf[ora_idx(o, i, j, k - 1)]; x1 = o->zc[k - 1]; }
                else { q1 = c_sdf[id]; x1 = o->zc[k]; }
                w_sdf[id] = ora_interp2(c_sdf[id], o->zc[k], q1, x1, o->zw[k]);
            }
}

/* interior (non-BOUNDARY) index ranges of each quantity, [lo, hi) */
static void q_ranges(const ora_ctx *o, int q, int *i0, int *i1, int *j0, int *j1, int *k0, int *k1) {
    *i0 = (q == TB_Q_U) ? 1 : 0; *i1 = o->NX - 1;
    *j0 = (q == TB_Q_V) ? 1 : 0; *j1 = o->NY - 1;
    *k0 = (q == TB_Q_W) ? 1 : 0; *k1 = o->NZ - 1;
}

/* Grid_tag_q_box_boundary_nodes (Grid.c:1617-1685),
 * Grid_tag_q_nodes_using_surface (Grid.c:1424-1612),
 * Grid_tag_q_buffer_solid_nodes (Grid.c:2024-2162, width 2, u/v/w only). */
static void classify(ora_ctx *o, int q) {
    int NX = o->NX, NY = o->NY, NZ = o->NZ;
    int NI = NX - 1, NJ = NY - 1, NK = NZ - 1;
    uint8_t *s = o->st[q];
    memset(s, 0, o->NT);
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++) {
            s[ora_idx(o, NX - 1, j, k)] = TB_BOUNDARY;
            if (q == TB_Q_U) s[ora_idx(o, 0, j, k)] = TB_BOUNDARY;
        }
    for (int k = 0; k < NZ; k++)
        for (int i = 0; i < NX; i++) {
            s[ora_idx(o, i, NY - 1, k)] = TB_BOUNDARY;
            if (q == TB_Q_V) s[ora_idx(o, i, 0, k)] = TB_BOUNDARY;
        }
    for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
            s[ora_idx(o, i, j, NZ - 1)] = TB_BOUNDARY;
            if (q == TB_Q_W) s[ora_idx(o, i, j, 0)] = TB_BOUNDARY;
        }
    int i0, i1, j0, j1, k0, k1;
    q_ranges(o, q, &i0, &i1, &j0, &j1, &k0, &k1);
    if (!o->has_geometry) {
        for (int k = k0; k < k1; k++)
            for (int j = j0; j < j1; j++)
                for (int i = i0; i < i1; i++) s[ora_idx(o, i, j, k)] = TB_FLUID;
        return;
    }
    const double *sdf = o->f[TB_F_SDF_U + q];
    for (int k = k0; k < k1; k++)
        for (int j = j0; j < j1; j++)
            for (int i = i0; i < i1; i++) {
                double s_p = sdf[ora_idx(o, i, j, k)];
                if (s_p >= 0.0) {
                    s[ora_idx(o, i, j, k)] = TB_FLUID;
                } else {
                    double s_e = sdf[ora_idx(o, ORA_MIN(i + 1, NI - 1), j, k)];
                    double s_w = sdf[ora_idx(o, ORA_MAX(i - 1, 0), j, k)];
                    double s_n = sdf[ora_idx(o, i, ORA_MIN(j + 1, NJ - 1), k)];
                    double s_s = sdf[ora_idx(o, i, ORA_MAX(j - 1, 0), k)];
                    double s_f = sdf[ora_idx(o, i, j, ORA_MIN(k + 1, NK - 1))];
                    double s_b = (double)sgn(sdf[ora_idx(o, i, j, ORA_MAX(k - 1, 0))]);
                    if (((s_p * s_e) < 0.0) || ((s_p * s_w) < 0.0) || ((s_p * s_n) < 0.0) ||
                        ((s_p * s_s) < 0.0) || ((s_p * s_b) < 0.0) || ((s_p * s_f) < 0.0))
                        s[ora_idx(o, i, j, k)] = TB_IMMERSED;
                    else
                        s[ora_idx(o, i, j, k)] = TB_SOLID;
                }
            }
    if (q == TB_Q_C) return;
    /* BUFFER_SOLID: reads the pre-pass (ghosted local copy), writes global */
    uint8_t *pre = (uint8_t *)malloc(o->NT);
    memcpy(pre, s, o->NT);
    const int width = 2;
    for (int k = k0; k < k1; k++)
        for (int j = j0; j < j1; j++)
            for (int i = i0; i < i1; i++) {
                if (pre[ora_idx(o, i, j, k)] != TB_SOLID) continue;
                int done = 0;
                for (int kz = -width; kz <= width && !done; kz++)
                    for (int ky = -width; ky <= width && !done; ky++)
                        for (int kx = -width; kx <= width && !done; kx++) {
                            int in = i + kx, jn = j + ky, kn = k + kz;
                            int stn = (in < 0 || in > NX - 1 || jn < 0 || jn > NY - 1 || kn < 0 ||
                                       kn > NZ - 1)
                                          ? TB_OUTSIDE
                                          : pre[ora_idx(o, in, jn, kn)];
                            if (stn == TB_IMMERSED) {
                                s[ora_idx(o, i, j, k)] = TB_BUFFER_SOLID;
                                done = 1;
                            }
                        }
            }
    free(pre);
}

/* ---------------------------------------------------------------------- */
/* Immersed-boundary nodes (Immersed.c)                                    */

static double point_dist(const double *a, const double *b) {
    /* MyMath_get_point_point_distance, MyMath.c:105-111 */
    double d2 = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    return sqrt(d2);
}

/* MyMath_binary_search (MyMath.c:384-398) */
static int binary_search(const double *data, double v, int lo, int hi) {
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (v >= data[mid]) {
            if (v < data[mid + 1]) return mid;
            lo = mid + 1;
        } else if (v < data[mid]) {
            hi = mid - 1;
        }
    }
    return -1;
}

/* Grid_get_x/y/z_index (Grid.c:1877-1967); note the z search uses >= */
static int get_index(const ora_ctx *o, const double *xq, double x, int dir) {
    int N = (dir == 0) ? o->NX - 1 : (dir == 1) ? o->NY - 1 : o->NZ - 1;
    if (x < xq[0]) return ORA_NOT_FOUND;
    if (dir == 2) { if (x >= xq[N - 1]) return ORA_GVG_INF; }
    else { if (x > xq[N - 1]) return ORA_GVG_INF; }
    return binary_search(xq, x, 0, N - 1);
}

void ora_gauss_jordan(int n, const double *a, double *ainv) {
    /* MyMath_GaussJordan_matrix_inv (MyMath.c:117-209), full pivoting */
    int index_col[8], index_row[8], index_piv[8];
    double A[8][8];
    int i, j, k, h, hh, in_row = -1, in_col = -1;
    for (j = 0; j < n; j++) for (i = 0; i < n; i++) A[j][i] = a[j * n + i];
    for (j = 0; j < n; j++) index_piv[j] = -1;
    for (i = 0; i < n; i++) {
        double mx = 0.0;
        for (j = 0; j < n; j++)
            if (index_piv[j] != 0)
                for (k = 0; k < n; k++)
                    if (index_piv[k] == -1 && fabs(A[j][k]) >= mx) {
                        mx = fabs(A[j][k]); in_row = j; in_col = k;
                    }
        ++(index_piv[in_col]);
        if (in_row != in_col)
            for (h = 0; h < n; h++) { double t = A[in_row][h]; A[in_row][h] = A[in_col][h]; A[in_col][h] = t; }
        index_row[i] = in_row;
        index_col[i] = in_col;
        double a_piv = 1.0 / A[in_col][in_col];
        A[in_col][in_col] = 1.0;
        for (h = 0; h < n; h++) A[in_col][h] *= a_piv;
        for (hh = 0; hh < n; hh++)
            if (hh != in_col) {
                double t = A[hh][in_col];
                A[hh][in_col] = 0.0;
                for (h = 0; h < n; h++) A[hh][h] -= A[in_col][h] * t;
            }
    }
    for (h = n - 1; h >= 0; h--)
        if (index_row[h] != index_col[h])
            for (k = 0; k < n; k++) {
                double t = A[k][index_row[h]];
                A[k][index_row[h]] = A[k][index_col[h]];
                A[k][index_col[h]] = t;
            }
    for (j = 0; j < n; j++) for (i = 0; i < n; i++) ainv[j * n + i] = A[j][i];
}

/* Immersed_find_box (Immersed.c:1210-1402) */
static void find_box(const ora_ctx *o, int q, tb_ib_node *nd, int *bi, int *bj, int *bk) {
    static const int dir_x[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    static const int dir_y[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int dir_z[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const double *xq, *yq, *zq;
    ora_q_coords(o, q, &xq, &yq, &zq);
    *bi = *bj = *bk = ORA_NOT_FOUND;
    double x = nd->image[0], y = nd->image[1], z = nd->image[2];
    const double tol_x = ORA_MAX(0.005 * o->dx_c[nd->i], 1e-12);
    const double tol_y = ORA_MAX(0.005 * o->dy_c[nd->j], 1e-12);
    const double tol_z = ORA_MAX(0.005 * o->dz_c[nd->k], 1e-12);
    const double pert_x[8] = {+tol_x, -tol_x, +tol_x, -tol_x, +tol_x, -tol_x, +tol_x, -tol_x};
    const double pert_y[8] = {+tol_y, +tol_y, -tol_y, -tol_y, +tol_y, +tol_y, -tol_y, -tol_y};
    const double pert_z[8] = {+tol_z, +tol_z, +tol_z, +tol_z, -tol_z, -tol_z, -tol_z, -tol_z};
    int m = 0, found = 0, i_ig = 0, j_ig = 0, k_ig = 0;
    while (!found && m < 8) {
        double x_p = x + pert_x[m], y_p = y + pert_y[m], z_p = z + pert_z[m];
        i_ig = get_index(o, xq, x_p, 0);
        j_ig = get_index(o, yq, y_p, 1);
        k_ig = get_index(o, zq, z_p, 2);
        int good = 1;
        for (int g = 0; g < 8; g++) {
            int stn = ora_status(o, q, i_ig + dir_x[g], j_ig + dir_y[g], k_ig + dir_z[g]);
            if (stn != TB_FLUID && stn != TB_IMMERSED) { good = 0; break; }
        }
        found = good;
        if (found) {
            if (!((x >= xq[i_ig]) && (x <= xq[i_ig + 1]) && (y >= yq[j_ig]) && (y <= yq[j_ig + 1]) &&
                  (z >= zq[k_ig]) && (z <= zq[k_ig + 1])))
                found = 0;
        }
        m++;
    }
    if (found) {
        nd->n_fluid = 8;
        *bi = i_ig; *bj = j_ig; *bk = k_ig;
        return;
    }
    /* fallback: nearest FLUID corner of the cell holding the image point */
    double D_min = 1e12;
    int g_found = ORA_NOT_FOUND;
    i_ig = get_index(o, xq, x, 0);
    j_ig = get_index(o, yq, y, 1);
    k_ig = get_index(o, zq, z, 2);
    i_ig = ORA_MAX(i_ig, 0); j_ig = ORA_MAX(j_ig, 0); k_ig = ORA_MAX(k_ig, 0);
    i_ig = ORA_MIN(i_ig, o->NX - 2); j_ig = ORA_MIN(j_ig, o->NY - 2); k_ig = ORA_MIN(k_ig, o->NZ - 2);
    /* (the NOT_FOUND / GVG_INF test after clamping, Immersed.c:1358-1368, can
     * no longer fire) */
    double p[3] = {x, y, z};
    for (int g = 0; g < 8; g++) {
        int i_ = i_ig + dir_x[g], j_ = j_ig + dir_y[g], k_ = k_ig + dir_z[g];
        if (ora_status(o, q, i_, j_, k_) == TB_FLUID) {
            double pp[3] = {xq[i_], yq[j_], zq[k_]};
            double d = point_dist(p, pp);
            if (d < D_min) { D_min = d; g_found = g; }
            found = 1;
        }
    }
    if (found) {
        nd->n_fluid = 1;
        *bi = i_ig + dir_x[g_found]; *bj = j_ig + dir_y[g_found]; *bk = k_ig + dir_z[g_found];
    } else {
        nd->n_fluid = 0;
    }
}

typedef struct { double p[8][3]; } ib_points;

/* Immersed_compute_interpolation_coef (Immersed.c:122-374) */
static void ib_coef_full(tb_ib_node *nd, const ib_points *fpnt, int bc) {
    double weight[8] = {0};
    int n_fluid = nd->n_fluid;
    double df = point_dist(nd->image, nd->control);
    double ds = point_dist(nd->im, nd->control);
    const double tol = 1e-10;
    double coef = ((fabs(ds) <= tol) && (fabs(df) <= tol)) ? 1.0 : ds / df;
    if (n_fluid == 8) {
        double x7 = fpnt->p[7][0], y7 = fpnt->p[7][1], z7 = fpnt->p[7][2];
        double x0 = fpnt->p[0][0], y0 = fpnt->p[0][1], z0 = fpnt->p[0][2];
        double alpha = (x7 - nd->image[0]) / (x7 - x0);
        double beta = (y7 - nd->image[1]) / (y7 - y0);
        double gamma = (z7 - nd->image[2]) / (z7 - z0);
        weight[0] = alpha * beta * gamma;
        weight[1] = (1.0 - alpha) * beta * gamma;
        weight[2] = alpha * (1.0 - beta) * gamma;
        weight[3] = (1.0 - alpha) * (1.0 - beta) * gamma;
        weight[4] = alpha * beta * (1.0 - gamma);
        weight[5] = (1.0 - alpha) * beta * (1.0 - gamma);
        weight[6] = alpha * (1.0 - beta) * (1.0 - gamma);
        weight[7] = (1.0 - alpha) * (1.0 - beta) * (1.0 - gamma);
    } else if (n_fluid == 1) {
        weight[0] = 1.0;
    }
    for (int g = 0; g < 8; g++) nd->coef[g] = 0.0;
    if (bc == ORA_DIRICHLET) {
        nd->boundary_coef = 1.0 + coef;
        for (int g = 0; g < n_fluid; g++) nd->coef[g] = weight[g] * coef * (+1.0);
    } else {
        if (n_fluid == 1) {
            nd->boundary_coef = -(ds + df);
            nd->coef[0] = weight[0] * (-1.0);
        } else {
            double A[64], Ainv[64];
            memset(A, 0, sizeof(A));
            int self_included = 0, r = 0;
            for (int g = 0; g < n_fluid; g++) {
                if (nd->fi[g] == nd->i && nd->fj[g] == nd->j && nd->fk[g] == nd->k) {
                    self_included = 1;
                } else {
                    double x = fpnt->p[g][0] - nd->image[0];
                    double y = fpnt->p[g][1] - nd->image[1];
                    double z = fpnt->p[g][2] - nd->image[2];
                    A[r * 8 + 0] = x * y * z;
                    A[r * 8 + 1] = x * y;
                    A[r * 8 + 2] = x * z;
                    A[r * 8 + 3] = y * z;
                    A[r * 8 + 4] = x;
                    A[r * 8 + 5] = y;
                    A[r * 8 + 6] = z;
                    A[r * 8 + 7] = 1.0;
                    r++;
                }
            }
            if (self_included) {
                double xc = nd->control[0] - nd->image[0];
                double yc = nd->control[1] - nd->image[1];
                double zc = nd->control[2] - nd->image[2];
                double nx = nd->n[0], ny = nd->n[1], nz = nd->n[2];
                A[r * 8 + 0] = nx * yc * zc + ny * xc * zc + nz * xc * yc;
                A[r * 8 + 1] = yc * nx + xc * ny;
                A[r * 8 + 2] = zc * nx + xc * nz;
                A[r * 8 + 3] = zc * ny + yc * nz;
                A[r * 8 + 4] = nx;
                A[r * 8 + 5] = ny;
                A[r * 8 + 6] = nz;
                A[r * 8 + 7] = 0.0;
            }
            ora_gauss_jordan(8, A, Ainv);
            double rhs = self_included ? Ainv[7 * 8 + 7] : 0.0;
            nd->boundary_coef = -(ds + df) + rhs;
            r = 0;
            for (int g = 0; g < n_fluid; g++) {
                if (nd->fi[g] == nd->i && nd->fj[g] == nd->j && nd->fk[g] == nd->k) {
                    nd->coef[g] = 0.0;
                } else {
                    nd->coef[g] = -Ainv[7 * 8 + r];
                    r++;
                }
            }
        }
    }
    for (int g = 0; g < n_fluid; g++)
        if (isnan(nd->coef[g]) || isinf(nd->coef[g])) { nd->coef[g] = 0.0; nd->boundary_coef = 0.0; }
}

/* Immersed_setup_q_immersed_nodes (Immersed.c:546-617) */
static void ib_setup(ora_ctx *o, int q) {
    static const int dir_x[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    static const int dir_y[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int dir_z[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    int N = 0;
    for (size_t id = 0; id < o->NT; id++) if (o->st[q][id] == TB_IMMERSED) N++;
    free(o->ib[q]);
    o->nib[q] = N;
    o->ib[q] = (tb_ib_node *)calloc((size_t)(N > 0 ? N : 1), sizeof(tb_ib_node));
    const double *xq, *yq, *zq;
    ora_q_coords(o, q, &xq, &yq, &zq);
    const double *sdf = o->f[TB_F_SDF_U + q];
    int NI = o->NX - 1, NJ = o->NY - 1, NK = o->NZ - 1;
    if (q == TB_Q_U) NI++;
    if (q == TB_Q_V) NJ++;
    if (q == TB_Q_W) NK++;
    int g = 0;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                if (o->st[q][ora_idx(o, i, j, k)] != TB_IMMERSED) continue;
                tb_ib_node *nd = &o->ib[q][g++];
                nd->i = i; nd->j = j; nd->k = k;
                nd->im[0] = xq[i]; nd->im[1] = yq[j]; nd->im[2] = zq[k];
            }
#pragma omp parallel for schedule(dynamic, 64)
    for (int n = 0; n < N; n++) {
        tb_ib_node *nd = &o->ib[q][n];
        int i = nd->i, j = nd->j, k = nd->k;
        /* Surface_compute_normal (Surface.c:547-619) */
        int iw = ORA_MAX(i - 1, 0), ie = ORA_MIN(i + 1, NI - 1);
        int is = ORA_MAX(j - 1, 0), in = ORA_MIN(j + 1, NJ - 1);
        int ib = ORA_MAX(k - 1, 0), iff = ORA_MIN(k + 1, NK - 1);
        double Delta_x = xq[ie] - xq[iw];
        double Delta_y = yq[in] - yq[is];
        double Delta_z = zq[iff] - zq[ib];
        double Dsx = (sdf[ora_idx(o, ie, j, k)] - sdf[ora_idx(o, iw, j, k)]) / Delta_x;
        double Dsy = (sdf[ora_idx(o, i, in, k)] - sdf[ora_idx(o, i, is, k)]) / Delta_y;
        double Dsz = (sdf[ora_idx(o, i, j, iff)] - sdf[ora_idx(o, i, j, ib)]) / Delta_z;
        double a_norm = 1.0 / sqrt(Dsx * Dsx + Dsy * Dsy + Dsz * Dsz);
        nd->n[0] = Dsx * a_norm;
        nd->n[1] = Dsy * a_norm;
        nd->n[2] = Dsz * a_norm;
        /* Immersed_find_control_point / Surface_find_Dn_to_surface
         * (Immersed.c:891-951, Surface.c:475-542): same differences */
        double D_sdf_dn = nd->n[0] * Dsx + nd->n[1] * Dsy + nd->n[2] * Dsz;
        if ((fabs(D_sdf_dn) < 1.0e-2) || isnan(D_sdf_dn)) D_sdf_dn = 1.0e-2;
        double Dn = sdf[ora_idx(o, i, j, k)] / D_sdf_dn;
        nd->control[0] = nd->im[0] - nd->n[0] * Dn;
        nd->control[1] = nd->im[1] - nd->n[1] * Dn;
        nd->control[2] = nd->im[2] - nd->n[2] * Dn;
        /* Immersed_find_image_point (Immersed.c:1063-1149) */
        double ds = point_dist(nd->im, nd->control);
        double D = 2.0 * ds;
        nd->image[0] = nd->im[0] + nd->n[0] * D;
        nd->image[1] = nd->im[1] + nd->n[1] * D;
        nd->image[2] = nd->im[2] + nd->n[2] * D;
        /* Immersed_assign_q_neighboring_fluid_nodes (Immersed.c:697-807) */
        int bi, bj, bk;
        find_box(o, q, nd, &bi, &bj, &bk);
        ib_points pts;
        memset(&pts, 0, sizeof(pts));
        if (bi != ORA_NOT_FOUND) {
            if (nd->n_fluid == 8) {
                for (int m = 0; m < 8; m++) {
                    int i_f = bi + dir_x[m], j_f = bj + dir_y[m], k_f = bk + dir_z[m];
                    nd->fi[m] = i_f; nd->fj[m] = j_f; nd->fk[m] = k_f;
                    pts.p[m][0] = xq[i_f]; pts.p[m][1] = yq[j_f]; pts.p[m][2] = zq[k_f];
                }
            } else if (nd->n_fluid == 1) {
                nd->fi[0] = bi; nd->fj[0] = bj; nd->fk[0] = bk;
                pts.p[0][0] = xq[bi]; pts.p[0][1] = yq[bj]; pts.p[0][2] = zq[bk];
            }
        }
        ib_coef_full(nd, &pts, (q == TB_Q_C) ? ORA_NEUMANN : ORA_DIRICHLET);
    }
}

int ora_setup_geometry(ora_ctx *o, const double *h_xz) {
    o->has_geometry = (h_xz != NULL);
    o->reinit_its = 0;
    if (o->has_geometry) {
        /* Grid_describe_interface (Grid.c:712-1121): interface = h + offset;
         * Surface_initialize (Surface.c:96-192): guess y - h */
        double offset = o->cfg.bottom_offset ? o->dy_c[0] * 2.001 : 0.0;
        double *p0 = dalloc(o->NT);
        for (int k = 0; k < o->NZ; k++)
            for (int j = 0; j < o->NY; j++)
                for (int i = 0; i < o->NX; i++) {
                    double h = h_xz[(size_t)k * o->NX + i] + offset;
                    p0[ora_idx(o, i, j, k)] = o->yc[j] - h;
                }
        /* Surface_find_exact_sdf (Surface.c:197-210): 500 iterations max */
        o->reinit_its = levelset_reinit(o, p0, o->f[TB_F_SDF_C], 500);
        free(p0);
        find_q_sdf(o);
    } else {
        for (int q = 0; q < 4; q++)
            for (size_t id = 0; id < o->NT; id++) o->f[TB_F_SDF_U + q][id] = 1.0;
    }
    for (int q = 0; q < 4; q++) classify(o, q);
    for (int q = 0; q < 4; q++) ib_setup(o, q);
    for (int q = 0; q < 4; q++) o->helm_dt[q] = 0.0;  /* matrices depend on flags */
    return 0;
}

/* Install flags and an IB table computed elsewhere (for timing the per-step
 * path at sizes where the setup chain would dominate the CPU sample). */
int ora_set_geometry(ora_ctx *o, int q, const uint8_t *flags, const tb_ib_node *ib, int n) {
    if (q < 0 || q > 3 || n < 0) return -1;
    memcpy(o->st[q], flags, o->NT);
    free(o->ib[q]);
    o->ib[q] = (tb_ib_node *)calloc((size_t)(n > 0 ? n : 1), sizeof(tb_ib_node));
    if (n > 0) memcpy(o->ib[q], ib, sizeof(tb_ib_node) * (size_t)n);
    o->nib[q] = n;
    o->helm_dt[q] = 0.0;
    o->has_geometry = 1;
    return 0;
}

int ora_reinit_iterations(ora_ctx *o) { return o->reinit_its; }

int ora_get_flags(ora_ctx *o, int q, uint8_t *out) {
    if (q < 0 || q > 3) return -1;
    memcpy(out, o->st[q], o->NT);
    return 0;
}

int ora_ib_count(ora_ctx *o, int q) { return (q < 0 || q > 3) ? -1 : o->nib[q]; }

int ora_get_ib(ora_ctx *o, int q, tb_ib_node *out, int cap) {
    if (q < 0 || q > 3) return -1;
    int n = ORA_MIN(cap, o->nib[q]);
    memcpy(out, o->ib[q], sizeof(tb_ib_node) * (size_t)n);
    return n;
}
