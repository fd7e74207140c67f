/*
 * ora_flow.c -- oracle restatement of the per-step physics: ENO convection
 * with wall ghosts, velocity transposes, right-hand sides, the Helmholtz
 * and Poisson operators (assembled row by row as the reference does),
 * BiCGStab, PCG, projection, divergence and the SSP-RK2 step.
 * TEST INFRASTRUCTURE ONLY (see tb_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

static double *dalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }
static void build_pois(ora_ctx *o);

ora_ctx *ora_create(const tb_config *cfg) {
    if (!cfg || cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2) return NULL;
    ora_ctx *o = (ora_ctx *)calloc(1, sizeof(ora_ctx));
    o->cfg = *cfg;
    o->ncx = cfg->nx; o->ncy = cfg->ny; o->ncz = cfg->nz;
    o->NX = cfg->nx + 1; o->NY = cfg->ny + 1; o->NZ = cfg->nz + 1;
    o->NT = (size_t)o->NX * o->NY * o->NZ;
    ora_grid_init(o);
    for (int q = 0; q < 4; q++) { o->st[q] = (uint8_t *)calloc(o->NT, 1); o->old[q] = dalloc(o->NT); }
    for (int f = 0; f < TB_F_COUNT; f++) o->f[f] = dalloc(o->NT);
    o->tr1 = dalloc(o->NT); o->tr2 = dalloc(o->NT);
    o->ubc = dalloc(o->NT); o->vbc = dalloc(o->NT); o->wbc = dalloc(o->NT);
    o->nconc = 1;
    o->cs[0].c = o->f[TB_F_C]; o->cs[0].conv = o->f[TB_F_CONV_C];
    o->cs[0].rhs = o->f[TB_F_RHS_C]; o->cs[0].old = o->old[3];
    o->cs[0].p.type = cfg->conc_type; o->cs[0].p.pe = cfg->pe; o->cs[0].p.v_settle = cfg->v_settle;
    o->cs[0].p.alpha = 1.0; o->cs[0].p.phi_max = 0.0;
    ora_setup_geometry(o, NULL);
    return o;
}

void ora_destroy(ora_ctx *o) {
    if (!o) return;
    double *g[] = {o->xc, o->yc, o->zc, o->xu, o->yu, o->zu, o->xv, o->yv, o->zv, o->xw, o->yw, o->zw,
                   o->dx_c, o->dy_c, o->dz_c, o->mxc, o->myc, o->mzc, o->mxu, o->myv, o->mzw};
    for (size_t a = 0; a < sizeof(g) / sizeof(g[0]); a++) free(g[a]);
    o->f[TB_F_C] = o->cs[0].c; o->f[TB_F_CONV_C] = o->cs[0].conv;
    o->f[TB_F_RHS_C] = o->cs[0].rhs; o->old[3] = o->cs[0].old;
    for (int n = 1; n < TB_NCONC_MAX; n++) {
        free(o->cs[n].c); free(o->cs[n].conv); free(o->cs[n].rhs); free(o->cs[n].old);
    }
    for (int n = 0; n < TB_NCONC_MAX; n++) free(o->cs[n].dep);
    free(o->ctot);
    for (int q = 0; q < 4; q++) { free(o->st[q]); free(o->old[q]); free(o->ib[q]); ora_mat_free(&o->helm[q]); }
    for (int f = 0; f < TB_F_COUNT; f++) free(o->f[f]);
    free(o->tr1); free(o->tr2); free(o->ubc); free(o->vbc); free(o->wbc);
    ora_mat_free(&o->pois);
    ora_gmg_free(o);
    ora_open_free(o);
    free(o);
}

int ora_set_field(ora_ctx *o, int f, const double *in) {
    if (f < 0 || f >= TB_F_COUNT) return -1;
    memcpy(o->f[f], in, sizeof(double) * o->NT);
    return 0;
}

int ora_get_field(ora_ctx *o, int f, double *out) {
    if (f < 0 || f >= TB_F_COUNT) return -1;
    memcpy(out, o->f[f], sizeof(double) * o->NT);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* ENO_compute_convective_terms (ENO.c:103-245)                            */

void ora_eno_line(int order, int N, const double *D0, const double *Position,
                  const double *Velocity, double *VdQdX) {
    double *D1 = dalloc((size_t)N + 1), *D2 = dalloc((size_t)N + 1), *D3 = dalloc((size_t)N + 1);
    int i;
    for (i = 0; i < N - 1; i++) D1[i] = (D0[i + 1] - D0[i]) / (Position[i + 1] - Position[i]);
    if (order > 1)
        for (i = 1; i < N - 1; i++) D2[i] = (D1[i] - D1[i - 1]) / (Position[i + 1] - Position[i - 1]);
    if (order > 2)
        for (i = 1; i < N - 2; i++) D3[i] = (D2[i + 1] - D2[i]) / (Position[i + 2] - Position[i - 1]);
    for (i = order; i < N - order; i++) {
        int k = (Velocity[i] > 0.0) ? i - 1 : i;
        int kstar = k;
        double dQ = D1[k];
        if (order > 1) {
            double c;
            if (fabs(D2[k]) <= fabs(D2[k + 1])) { c = D2[k]; kstar = k - 1; }
            else { c = D2[k + 1]; kstar = k; }
            double x_i = Position[i], x_k = Position[k], x_k1 = Position[k + 1];
            dQ += c * ((x_i - x_k) + (x_i - x_k1));
        }
        if (order > 2) {
            double cstar = (fabs(D3[kstar]) <= fabs(D3[kstar + 1])) ? D3[kstar] : D3[kstar + 1];
            double x_i = Position[i], x_ks = Position[kstar], x_ks1 = Position[kstar + 1],
                   x_ks2 = Position[kstar + 2];
            dQ += cstar * ((x_i - x_ks) * (x_i - x_ks1) + (x_i - x_ks1) * (x_i - x_ks2) +
                           (x_i - x_ks) * (x_i - x_ks2));
        }
        VdQdX[i] = Velocity[i] * dQ;
    }
    free(D1); free(D2); free(D3);
}

/* Wall ghosts of one line (Velocity.c:2124-2801, Conc.c:1021-1220).
 * kind 0: normal component whose wall value sits on the grid (u-x, v-y, w-z);
 * kind 1: tangential / cell-centred line mirrored about the wall.
 * sgn_lo/sgn_hi: +1 copy first interior value, -1 negate it. */
static void ghost_fill(int order, int N, double *D0, double *P, int kind, double wall_lo,
                       double wall_hi, double sgn_lo, double sgn_hi, int u_x_quirk,
                       int hi_mode, double hi_scale) {
    int ex;
    for (ex = 0; ex < order; ex++) {
        int gi = ex;
        if (kind == 0) {
            int ii = 2 * order - ex - 2;
            if (ex == order - 1) { P[gi] = wall_lo; D0[gi] = 0.0; }
            else { P[gi] = 2.0 * wall_lo - P[ii]; D0[gi] = u_x_quirk ? -D0[order] : 0.0; }
        } else {
            int ii = 2 * order - ex - 1;
            P[gi] = 2.0 * wall_lo - P[ii];
            D0[gi] = (sgn_lo < 0.0) ? -D0[order] : +D0[order];
        }
    }
    for (ex = 0; ex < order; ex++) {
        int gi = N - ex - 1;
        if (kind == 0) {
            int ii = N - 2 * order + ex + 1;
            if (ex == order - 1) { P[gi] = wall_hi; D0[gi] = 0.0; }
            else { P[gi] = 2.0 * wall_hi - P[ii]; D0[gi] = 0.0; }
        } else {
            int ii = N - 2 * order + ex;
            P[gi] = 2.0 * wall_hi - P[ii];
            if (hi_mode == 1) D0[gi] = hi_scale * D0[N - order - 1];
            else D0[gi] = (sgn_hi < 0.0) ? -D0[N - order - 1] : +D0[N - order - 1];
        }
    }
}

/* Velocity_cell_center (Velocity.c:1476-1602); an open x face gives u
 * there from its plane (Velocity.c:1544-1558) */
static void cell_center(ora_ctx *o) {
    const double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    const double *uin = ora_in_plane(o, TB_Q_U), *uout = ora_out_plane(o, TB_Q_U);
    int NI = o->NX - 1, NJ = o->NY - 1, NK = o->NZ - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < NK; k++)
        for (int j = 0; j < NJ; j++)
            for (int i = 0; i < NI; i++) {
                size_t id = ora_idx(o, i, j, k);
                double u0 = (uin && i == 0) ? uin[ora_pidx(o, j, k)] : u[id];
                double u1 = (uout && i == NI - 1) ? uout[ora_pidx(o, j, k)] : u[id + 1];
                o->ubc[id] = ora_interp2(u0, o->xu[i], u1, o->xu[i + 1], o->xc[i]);
                o->vbc[id] = ora_interp2(v[id], o->yv[j], v[id + o->NX], o->yv[j + 1], o->yc[j]);
                size_t pl = (size_t)o->NX * o->NY;
                o->wbc[id] = ora_interp2(w[id], o->zw[k], w[id + pl], o->zw[k + 1], o->zc[k]);
            }
}

/* Velocity_transpose_velocities (Velocity.c:3341-3663); u at the v and w
 * nodes next to an open x face comes from its plane (Velocity.c:3491-3505,
 * 3589-3603) */
static void transpose(ora_ctx *o, int q) {
    const double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    const double *uin = ora_in_plane(o, TB_Q_U), *uout = ora_out_plane(o, TB_Q_U);
    double *t1 = o->tr1, *t2 = o->tr2;
    int NX = o->NX, NY = o->NY, NZ = o->NZ;
    size_t pl = (size_t)NX * NY;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                if (ora_status(o, q, i, j, k) == TB_BOUNDARY) continue;
                size_t id = ora_idx(o, i, j, k);
                double qq[4], X1[4], X2[4];
                if (q == TB_Q_U) {
                    qq[0] = v[id - 1]; qq[1] = v[id]; qq[2] = v[id + NX - 1]; qq[3] = v[id + NX];
                    X1[0] = o->xv[i - 1]; X1[1] = o->xv[i]; X1[2] = o->xv[i - 1]; X1[3] = o->xv[i];
                    X2[0] = o->yv[j]; X2[1] = o->yv[j]; X2[2] = o->yv[j + 1]; X2[3] = o->yv[j + 1];
                    t1[id] = ora_interp4(qq, X1, X2, o->xu[i], o->yu[j]);
                    qq[0] = w[id - 1]; qq[1] = w[id]; qq[2] = w[id + pl - 1]; qq[3] = w[id + pl];
                    X1[0] = o->xw[i - 1]; X1[1] = o->xw[i]; X1[2] = o->xw[i - 1]; X1[3] = o->xw[i];
                    X2[0] = o->zw[k]; X2[1] = o->zw[k]; X2[2] = o->zw[k + 1]; X2[3] = o->zw[k + 1];
                    t2[id] = ora_interp4(qq, X1, X2, o->xu[i], o->zu[k]);
                } else if (q == TB_Q_V) {
                    qq[0] = u[id - NX]; qq[2] = u[id];
                    qq[1] = u[id - NX + 1]; qq[3] = u[id + 1];
                    if (uin && i == 0) { qq[0] = uin[ora_pidx(o, j - 1, k)]; qq[2] = uin[ora_pidx(o, j, k)]; }
                    if (uout && i == NX - 2) { qq[1] = uout[ora_pidx(o, j - 1, k)]; qq[3] = uout[ora_pidx(o, j, k)]; }
                    X1[0] = o->xu[i]; X1[1] = o->xu[i + 1]; X1[2] = o->xu[i]; X1[3] = o->xu[i + 1];
                    X2[0] = o->yu[j - 1]; X2[1] = o->yu[j - 1]; X2[2] = o->yu[j]; X2[3] = o->yu[j];
                    t1[id] = ora_interp4(qq, X1, X2, o->xv[i], o->yv[j]);
                    qq[0] = w[id - NX]; qq[1] = w[id + pl - NX]; qq[2] = w[id]; qq[3] = w[id + pl];
                    X1[0] = o->zw[k]; X1[1] = o->zw[k + 1]; X1[2] = o->zw[k]; X1[3] = o->zw[k + 1];
                    X2[0] = o->yw[j - 1]; X2[1] = o->yw[j - 1]; X2[2] = o->yw[j]; X2[3] = o->yw[j];
                    t2[id] = ora_interp4(qq, X1, X2, o->zv[k], o->yv[j]);
                } else {
                    qq[0] = u[id - pl]; qq[2] = u[id];
                    qq[1] = u[id - pl + 1]; qq[3] = u[id + 1];
                    if (uin && i == 0) { qq[0] = uin[ora_pidx(o, j, k - 1)]; qq[2] = uin[ora_pidx(o, j, k)]; }
                    if (uout && i == NX - 2) { qq[1] = uout[ora_pidx(o, j, k - 1)]; qq[3] = uout[ora_pidx(o, j, k)]; }
                    X1[0] = o->xu[i]; X1[1] = o->xu[i + 1]; X1[2] = o->xu[i]; X1[3] = o->xu[i + 1];
                    X2[0] = o->zu[k - 1]; X2[1] = o->zu[k - 1]; X2[2] = o->zu[k]; X2[3] = o->zu[k];
                    t1[id] = ora_interp4(qq, X1, X2, o->xw[i], o->zw[k]);
                    qq[0] = v[id - pl]; qq[1] = v[id]; qq[2] = v[id - pl + NX]; qq[3] = v[id + NX];
                    X1[0] = o->zv[k - 1]; X1[1] = o->zv[k]; X1[2] = o->zv[k - 1]; X1[3] = o->zv[k];
                    X2[0] = o->yv[j]; X2[1] = o->yv[j]; X2[2] = o->yv[j + 1]; X2[3] = o->yv[j + 1];
                    t2[id] = ora_interp4(qq, X1, X2, o->zw[k], o->yw[j]);
                }
            }
}

/* One quantity, one direction: the reconstructed line driver (the per-line
 * gather of the missing gvg.c) around ENO_compute_convective_terms. Lines
 * span every non-BOUNDARY node between the two box walls. */
static void eno_direction(ora_ctx *o, int q, int d, const double *adv, double *conv, int first) {
    const double *D = o->f[TB_F_U + q];
    int order = o->cfg.eno_order;
    int lo[3] = {(q == TB_Q_U) ? 1 : 0, (q == TB_Q_V) ? 1 : 0, (q == TB_Q_W) ? 1 : 0};
    int hi[3] = {o->NX - 2, o->NY - 2, o->NZ - 2};
    const double *xq, *yq, *zq;
    ora_q_coords(o, q, &xq, &yq, &zq);
    const double *pos = (d == 0) ? xq : (d == 1) ? yq : zq;
    double wall_lo = (d == 0) ? o->xu[0] : (d == 1) ? o->yv[0] : o->zw[0];
    double wall_hi = (d == 0) ? o->xu[o->NX - 1] : (d == 1) ? o->yv[o->NY - 1] : o->zw[o->NZ - 1];
    int kind = (q == d) ? 0 : 1;
    int u_x_quirk = (q == TB_Q_U && d == 0);
    double s_lo = 1.0, s_hi = 1.0;
    int hi_mode = 0;
    double hi_scale = 1.0;
    if (q != TB_Q_C) {
        s_lo = o->cfg.wall_noslip[2 * d] ? -1.0 : 1.0;
        s_hi = o->cfg.wall_noslip[2 * d + 1] ? -1.0 : 1.0;
    } else if (d == 1 && o->cfg.conc_type == TB_PARTICLE && o->cfg.sedimentation) {
        double coef = 2.0 / (o->cfg.pe * o->dXi) * o->myv[o->NY - 1];
        double Vs = fabs(o->cfg.v_settle);
        hi_mode = 1;
        hi_scale = (coef - Vs) / (coef + Vs);
    }
    int n = hi[d] - lo[d] + 1;
    int N = n + 2 * order;
    size_t stride = (d == 0) ? 1 : (d == 1) ? (size_t)o->NX : (size_t)o->NX * o->NY;
    int a0 = (d == 0) ? 1 : 0, a1 = (d == 2) ? 1 : 2;   /* the two other axes */
    /* open x faces: every ghost of that side takes the plane value
     * (u: Velocity.c:2193-2200, 2227-2233; v, w: outflow only, 2442-2444,
     * 2669-2671; c: Conc.c:1099-1101, 1120-1122) */
    const double *gin = NULL, *gout = NULL;
    if (d == 0) {
        if (q == TB_Q_U || q == TB_Q_C) gin = ora_in_plane(o, q);
        gout = ora_out_plane(o, q);
    }
#pragma omp parallel
    {
        double *D0 = dalloc((size_t)N), *P = dalloc((size_t)N), *V = dalloc((size_t)N), *out = dalloc((size_t)N);
#pragma omp for schedule(static) collapse(2)
        for (int b = lo[a1]; b <= hi[a1]; b++)
            for (int a = lo[a0]; a <= hi[a0]; a++) {
                int ijk[3];
                ijk[a0] = a; ijk[a1] = b; ijk[d] = lo[d];
                size_t base = ora_idx(o, ijk[0], ijk[1], ijk[2]);
                for (int m = 0; m < n; m++) {
                    size_t id = base + (size_t)m * stride;
                    D0[order + m] = D[id];
                    P[order + m] = pos[lo[d] + m];
                    V[order + m] = adv[id];
                }
                ghost_fill(order, N, D0, P, kind, wall_lo, wall_hi, s_lo, s_hi, u_x_quirk, hi_mode,
                           hi_scale);
                if (gin)
                    for (int ex = 0; ex < order; ex++) D0[ex] = gin[ora_pidx(o, a, b)];
                if (gout)
                    for (int ex = 0; ex < order; ex++) D0[N - ex - 1] = gout[ora_pidx(o, a, b)];
                ora_eno_line(order, N, D0, P, V, out);
                for (int m = 0; m < n; m++) {
                    size_t id = base + (size_t)m * stride;
                    if (first) conv[id] = out[order + m];
                    else conv[id] += out[order + m];
                }
            }
        free(D0); free(P); free(V); free(out);
    }
}

int ora_eno_conv(ora_ctx *o, int q) {
    if (q < 0 || q > 3) return -1;
    double *conv = o->f[TB_F_CONV_U + q];
    memset(conv, 0, sizeof(double) * o->NT);
    if (q == TB_Q_C) {
        cell_center(o);
        const double *vadv = o->vbc;
        double *vp = NULL;
        if (o->cfg.conc_type == TB_PARTICLE) {
            /* Conc_set_particle_v_velocity (Conc.c:734-761): VecShift by
             * V_s0, or with hindered settling VecWAXPY with the field of
             * Conc_compute_particle_settling_speed (Conc.c:764-812): the
             * hindered speed on FLUID cells, 0 elsewhere */
            const double *cd = o->f[TB_F_C];
            double phi = o->cs[o->conc_sel].p.phi_max, vs = o->cfg.v_settle;
            vp = dalloc(o->NT);
            for (size_t id = 0; id < o->NT; id++) {
                if (!o->hindered) vp[id] = o->vbc[id] + vs;
                else {
                    double v_settl = (o->st[TB_Q_C][id] == TB_FLUID) ? ora_hindered_speed(cd[id], phi, vs) : 0.0;
                    vp[id] = o->vbc[id] + v_settl;
                }
            }
            vadv = vp;
        }
        eno_direction(o, q, 0, o->ubc, conv, 1);
        eno_direction(o, q, 1, vadv, conv, 0);
        eno_direction(o, q, 2, o->wbc, conv, 0);
        free(vp);
        return 0;
    }
    transpose(o, q);
    const double *adv[3];
    if (q == TB_Q_U) { adv[0] = o->f[TB_F_U]; adv[1] = o->tr1; adv[2] = o->tr2; }
    else if (q == TB_Q_V) { adv[0] = o->tr1; adv[1] = o->f[TB_F_V]; adv[2] = o->tr2; }
    else { adv[0] = o->tr1; adv[1] = o->tr2; adv[2] = o->f[TB_F_W]; }
    eno_direction(o, q, 0, adv[0], conv, 1);
    eno_direction(o, q, 1, adv[1], conv, 0);
    eno_direction(o, q, 2, adv[2], conv, 0);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* right-hand sides: Velocity.c:1607-2043, Conc.c:864-992                  */

int ora_set_rhs(ora_ctx *o, int q, double dt) {
    if (q < 0 || q > 3) return -1;
    /* open x faces: the plane value times the dropped coefficient
     * (u: Velocity.c:1691-1721; v: 1885-1897; w: 2011-2023; c: Conc.c:938-963) */
    const double *pin = (q == TB_Q_U || q == TB_Q_C) ? ora_in_plane(o, q) : NULL;
    const double *pout = ora_out_plane(o, q);
    double kappa = (q == TB_Q_C) ? o->cfg.pe : o->cfg.re;
    double Cx = -1.0 / (kappa * o->dEta * o->dEta);
    const double *data = o->f[TB_F_U + q], *conv = o->f[TB_F_CONV_U + q];
    const double *dp = (q < 3) ? o->f[TB_F_DPDX + q] : NULL;
    /* Velocity.c:1798-1815: c_total with several fields, else c[0] */
    const double *c = (o->nconc > 1) ? o->ctot : o->cs[0].c;
    double *rhs = o->f[TB_F_RHS_U + q];
    double a_dt = 1.0 / dt;
    double buoy = o->cfg.buoyancy;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                int s = o->st[q][id];
                double r = 0.0;
                if (q == TB_Q_C) {
                    if (s == TB_FLUID) {
                        r = data[id] * a_dt - conv[id];
                        if (pin && i == 0) {
                            double aw = -Cx * o->mxc[i] * o->mxu[i];
                            double value = aw * pin[ora_pidx(o, j, k)];
                            r += value;
                        }
                        if (pout && i == o->NX - 2) {
                            double ae = -Cx * o->mxc[i] * o->mxu[i];   /* metric_xu[i], Conc.c:957 */
                            double value = ae * pout[ora_pidx(o, j, k)];
                            r += value;
                        }
                    }
                } else if (s == TB_FLUID || s == TB_BUFFER_SOLID) {
                    r = data[id] * a_dt - conv[id] - dp[id];
                    if (q == TB_Q_V) {
                        double c_term = ora_interp2(c[id - o->NX], o->yc[j - 1], c[id], o->yc[j], o->yv[j]);
                        r = data[id] * a_dt - conv[id] - dp[id] - buoy * c_term;
                    }
                    if (q == TB_Q_U) {
                        if (pin && i == 1) {
                            double aw = -Cx * o->mxu[i] * o->mxc[i - 1];
                            double value = aw * pin[ora_pidx(o, j, k)];
                            r += value;
                        }
                        if (pout && i == o->NX - 2) {
                            double ae = -Cx * o->mxu[i] * o->mxc[i];
                            double value = ae * pout[ora_pidx(o, j, k)];
                            r += value;
                        }
                    } else if (pout && i == o->NX - 2) {
                        double ae = -Cx * o->mxc[i] * o->mxu[i + 1];
                        double value = ae * pout[ora_pidx(o, j, k)];
                        r += value;
                    }
                }
                rhs[id] = r;
            }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* sparse rows                                                              */

void ora_mat_free(ora_mat *m) {
    free(m->cnt); free(m->col); free(m->val);
    memset(m, 0, sizeof(*m));
}

static void mat_alloc(ora_mat *m, size_t n) {
    ora_mat_free(m);
    m->n = n;
    m->cnt = (int32_t *)calloc(n, sizeof(int32_t));
    m->col = (int64_t *)calloc(n * 9, sizeof(int64_t));
    m->val = (double *)calloc(n * 9, sizeof(double));
}

static void row_set(ora_mat *m, size_t row, int cnt, const int64_t *cols, const double *vals) {
    /* insertion sort by column (PETSc AIJ rows are sorted) */
    int64_t c[9];
    double v[9];
    for (int a = 0; a < cnt; a++) { c[a] = cols[a]; v[a] = vals[a]; }
    for (int a = 1; a < cnt; a++) {
        int64_t cc = c[a]; double vv = v[a]; int b = a - 1;
        while (b >= 0 && c[b] > cc) { c[b + 1] = c[b]; v[b + 1] = v[b]; b--; }
        c[b + 1] = cc; v[b + 1] = vv;
    }
    m->cnt[row] = cnt;
    for (int a = 0; a < cnt; a++) { m->col[row * 9 + a] = c[a]; m->val[row * 9 + a] = v[a]; }
}

static void row_add_diag(ora_mat *m, size_t row, double v) {
    for (int a = 0; a < m->cnt[row]; a++)
        if ((size_t)m->col[row * 9 + a] == row) { m->val[row * 9 + a] += v; return; }
}

void ora_mat_apply(const ora_mat *m, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < m->n; r++) {
        double s = 0.0;
        const int64_t *c = m->col + r * 9;
        const double *v = m->val + r * 9;
        for (int a = 0; a < m->cnt[r]; a++) s += v[a] * x[c[a]];
        y[r] = s;
    }
}

/* Velocity_u/v/w_form_LHS_matrix (Velocity.c:449-1471), Conc_form_LHS_matrix
 * (Conc.c:390-730), then Velocity_modify_diagonal / Conc_modify_diagonal
 * (Velocity.c:374-443, Conc.c:293-354) with dt_old = 0 (mod = 1/dt). */
static void build_helm(ora_ctx *o, int q, double dt) {
    ora_mat *m = &o->helm[q];
    mat_alloc(m, o->NT);
    double kappa = (q == TB_Q_C) ? o->cfg.pe : o->cfg.re;
    double Cx = -1.0 / (kappa * o->dEta * o->dEta);
    double Cy = -1.0 / (kappa * o->dXi * o->dXi);
    double Cz = -1.0 / (kappa * o->dPhi * o->dPhi);
    double ap_solid = -Cx * o->mxc[0] - Cy * o->myc[0] - Cz * o->mzc[0];
    const int *wall = o->cfg.wall_noslip;
    int NX = o->NX, NY = o->NY;
    size_t pl = (size_t)NX * NY;
    int vNI = NX - 1, uNJ = NY - 1, uNK = o->NZ - 1;
    double mod = 1.0 / dt;
    /* open x faces drop their wall rule: v, w at x = Lx (Velocity.c:989,
     * 1336), c at x = 0 and x = Lx (Conc.c:582-607) */
    int open_w = (q == TB_Q_C) && o->obc.inflow;
    int open_e = (q != TB_Q_U) && o->obc.outflow;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                int s = o->st[q][id];
                int64_t cols[9];
                double vals[9];
                int nnz = 0;
                int row7 = (q == TB_Q_C) ? (s == TB_FLUID) : (s == TB_FLUID || s == TB_BUFFER_SOLID);
                if (row7) {
                    double mpx, mpy, mpz, mex, mwx, mny, msy, mbz, mfz, ap;
                    if (q == TB_Q_U) {
                        mpx = o->mxu[i]; mpy = o->myc[j]; mpz = o->mzc[k]; mex = o->mxc[i]; mwx = o->mxc[i - 1];
                        mny = o->myv[j + 1]; msy = o->myv[j]; mbz = o->mzw[k]; mfz = o->mzw[k + 1];
                        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mfz + mbz);
                    } else if (q == TB_Q_V) {
                        mpx = o->mxc[i]; mpy = o->myv[j]; mpz = o->mzc[k]; mex = o->mxu[i + 1]; mwx = o->mxu[i];
                        mny = o->myc[j]; msy = o->myc[j - 1]; mbz = o->mzw[k]; mfz = o->mzw[k + 1];
                        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
                    } else if (q == TB_Q_W) {
                        mpx = o->mxc[i]; mpy = o->myc[j]; mpz = o->mzw[k]; mex = o->mxu[i + 1]; mwx = o->mxu[i];
                        mny = o->myv[j + 1]; msy = o->myv[j]; mbz = o->mzc[k - 1]; mfz = o->mzc[k];
                        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
                    } else {
                        mpx = o->mxc[i]; mpy = o->myc[j]; mpz = o->mzc[k]; mex = o->mxu[i + 1]; mwx = o->mxu[i];
                        mny = o->myv[j + 1]; msy = o->myv[j]; mbz = o->mzw[k]; mfz = o->mzw[k + 1];
                        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
                    }
                    double aw = Cx * mpx * mwx, ae = Cx * mpx * mex;
                    double an = Cy * mpy * mny, as = Cy * mpy * msy;
                    double ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
#define ADD(v, c) do { vals[nnz] = (v); cols[nnz] = (int64_t)(c); nnz++; } while (0)
#define WALL(a, w) do { if (wall[w]) ap -= (a); else ap += (a); } while (0)
                    if (q == TB_Q_U) {
                        if (k == 0) WALL(ab, 4); else ADD(ab, id - pl);
                        if (j == 0) WALL(as, 2); else ADD(as, id - NX);
                        if (ora_status(o, q, i - 1, j, k) != TB_BOUNDARY) ADD(aw, id - 1);
                        if (ora_status(o, q, i + 1, j, k) != TB_BOUNDARY) ADD(ae, id + 1);
                        if (j == uNJ - 1) WALL(an, 3); else ADD(an, id + NX);
                        if (k == uNK - 1) WALL(af, 5); else ADD(af, id + pl);
                    } else if (q == TB_Q_V) {
                        if (k == 0) WALL(ab, 4); else ADD(ab, id - pl);
                        if (ora_status(o, q, i, j - 1, k) != TB_BOUNDARY) ADD(as, id - NX);
                        if (i == 0) WALL(aw, 0); else ADD(aw, id - 1);
                        if (i == vNI - 1) { if (!open_e) WALL(ae, 1); } else if (i < vNI - 1) ADD(ae, id + 1);
                        if (ora_status(o, q, i, j + 1, k) != TB_BOUNDARY) ADD(an, id + NX);
                        if (k == uNK - 1) WALL(af, 5); else ADD(af, id + pl);
                    } else if (q == TB_Q_W) {
                        if (ora_status(o, q, i, j, k - 1) != TB_BOUNDARY) ADD(ab, id - pl);
                        if (j == 0) WALL(as, 2); else ADD(as, id - NX);
                        if (i == 0) WALL(aw, 0); else ADD(aw, id - 1);
                        if (i == vNI - 1) { if (!open_e) WALL(ae, 1); } else if (i < vNI - 1) ADD(ae, id + 1);
                        if (j == uNJ - 1) WALL(an, 3); else ADD(an, id + NX);
                        if (ora_status(o, q, i, j, k + 1) != TB_BOUNDARY) ADD(af, id + pl);
                    } else {
                        int cNI = NX - 1, cNJ = NY - 1, cNK = o->NZ - 1;
                        if (k > 0) ADD(ab, id - pl); else ap += ab;
                        if (j > 0) ADD(as, id - NX); else ap += as;
                        if (i > 0) ADD(aw, id - 1); else if (!open_w) ap += aw;
                        if (i < cNI - 1) ADD(ae, id + 1); else if (!(open_e && i == cNI - 1)) ap += ae;
                        if (j < cNJ - 1) ADD(an, id + NX);
                        else {
                            double coef = an;
                            if (o->cfg.conc_type == TB_PARTICLE) {
                                double b = 2.0 / (o->cfg.pe * o->dXi) * mny;
                                coef = an * (b - fabs(o->cfg.v_settle)) / (b + fabs(o->cfg.v_settle));
                            }
                            ap += coef;
                        }
                        if (k < cNK - 1) ADD(af, id + pl); else ap += af;
                    }
                    ADD(ap, id);
                    ap = 0.0;
#undef WALL
                } else if (s == TB_IMMERSED) {
                    /* IB row scaled by ap_solid (Velocity.c:718-762, Conc.c:668-708) */
                    continue;   /* filled below, sequentially over the IB list */
                } else {
                    ADD(ap_solid, id);
                }
                row_set(m, id, nnz, cols, vals);
                if (s != TB_IMMERSED) row_add_diag(m, id, mod);
            }
    for (int n = 0; n < o->nib[q]; n++) {
        const tb_ib_node *nd = &o->ib[q][n];
        size_t id = ora_idx(o, nd->i, nd->j, nd->k);
        int64_t cols[9];
        double vals[9];
        int nnz = 0;
        ADD(1.0, id);
        for (int g = 0; g < nd->n_fluid; g++) {
            if (nd->fi[g] == nd->i && nd->fj[g] == nd->j && nd->fk[g] == nd->k) vals[0] += nd->coef[g];
            else if (fabs(nd->coef[g]) >= ORA_DROP_TOL) ADD(nd->coef[g], ora_idx(o, nd->fi[g], nd->fj[g], nd->fk[g]));
        }
        for (int s2 = 0; s2 < nnz; s2++) vals[s2] *= ap_solid;
        row_set(m, id, nnz, cols, vals);
    }
#undef ADD
    o->helm_dt[q] = dt;
}

static void ensure_helm(ora_ctx *o, int q, double dt) {
    if (o->helm_dt[q] != dt || o->helm[q].n != o->NT) build_helm(o, q, dt);
}

/* Copy out the assembled rows (9 slots per row, columns sorted), e.g. for a
 * scipy direct solve in the tests. */
int ora_helmholtz_matrix(ora_ctx *o, int q, double dt, int32_t *cnt, int64_t *col, double *val) {
    if (q < 0 || q > 3) return -1;
    ensure_helm(o, q, dt);
    memcpy(cnt, o->helm[q].cnt, sizeof(int32_t) * o->NT);
    memcpy(col, o->helm[q].col, sizeof(int64_t) * o->NT * 9);
    memcpy(val, o->helm[q].val, sizeof(double) * o->NT * 9);
    return 0;
}

int ora_poisson_matrix(ora_ctx *o, int32_t *cnt, int64_t *col, double *val) {
    if (!o->pois_built) build_pois(o);
    memcpy(cnt, o->pois.cnt, sizeof(int32_t) * o->NT);
    memcpy(col, o->pois.col, sizeof(int64_t) * o->NT * 9);
    memcpy(val, o->pois.val, sizeof(double) * o->NT * 9);
    return 0;
}

int ora_helmholtz_apply(ora_ctx *o, int q, double dt, const double *x, double *y) {
    if (q < 0 || q > 3) return -1;
    ensure_helm(o, q, dt);
    ora_mat_apply(&o->helm[q], x, y);
    return 0;
}

/* Gauss-Seidel value of row r of the assembled rows: (b_r - sum_{c != r}
 * a_rc x_c) / a_rr, off-diagonal terms in ascending column order */
static double gs_row(const ora_mat *m, size_t r, const double *b, const double *x) {
    double s = 0.0, dg = 1.0;
    for (int e = 0; e < m->cnt[r]; e++) {
        int64_t c = m->col[r * 9 + e];
        double v = m->val[r * 9 + e];
        if ((size_t)c == r) dg = v;
        else s += v * x[c];
    }
    return (b[r] - s) / dg;
}

/* Red-black Gauss-Seidel sweep pairs on the assembled Helmholtz rows (the
 * smoother of the device's Helmholtz solve; the reference solves with KSP):
 * red = (i + j + k) even first, each colour in place, IMMERSED rows skipped
 * by both colours and then updated once per pair by Jacobi. */
int ora_helmholtz_gs_sweeps(ora_ctx *o, int q, double dt, const double *b, double *x, int pairs) {
    if (q < 0 || q > 3 || pairs < 0) return -1;
    ensure_helm(o, q, dt);
    const ora_mat *m = &o->helm[q];
    const uint8_t *st = o->st[q];
    double *tmp = dalloc(o->nib[q] > 0 ? (size_t)o->nib[q] : 1);
    for (int p = 0; p < pairs; p++) {
        for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(static)
            for (int k = 0; k < o->NZ; k++)
                for (int j = 0; j < o->NY; j++)
                    for (int i = (j + k + color) & 1; i < o->NX; i += 2) {
                        size_t id = ora_idx(o, i, j, k);
                        if (st[id] != TB_IMMERSED) x[id] = gs_row(m, id, b, x);
                    }
        }
        for (int n = 0; n < o->nib[q]; n++) {
            const tb_ib_node *nd = &o->ib[q][n];
            tmp[n] = gs_row(m, ora_idx(o, nd->i, nd->j, nd->k), b, x);
        }
        for (int n = 0; n < o->nib[q]; n++) {
            const tb_ib_node *nd = &o->ib[q][n];
            x[ora_idx(o, nd->i, nd->j, nd->k)] = tmp[n];
        }
    }
    free(tmp);
    return 0;
}

/* a . b as DOT_BLOCKS block sums added in block order: the same value
 * whatever the number of threads or their finishing order (an OpenMP
 * reduction combines the threads' partials in arrival order) */
#define DOT_BLOCKS 256
static double dot(const double *a, const double *b, size_t n) {
    double part[DOT_BLOCKS];
    size_t per = (n + DOT_BLOCKS - 1) / DOT_BLOCKS;
#pragma omp parallel for schedule(static)
    for (int blk = 0; blk < DOT_BLOCKS; blk++) {
        size_t i0 = (size_t)blk * per, i1 = i0 + per < n ? i0 + per : n;
        double s = 0.0;
        for (size_t i = i0; i < i1; i++) s += a[i] * b[i];
        part[blk] = s;
    }
    double s = 0.0;
    for (int blk = 0; blk < DOT_BLOCKS; blk++) s += part[blk];
    return s;
}

/* BiCGStab without preconditioner (the KSPIBCGS + PCNONE of Solver.c:269-322
 * in its textbook form), nonzero initial guess = current field. */
int ora_helmholtz_solve(ora_ctx *o, int q, double dt, double rtol, int maxit, tb_solve_stats *st) {
    if (q < 0 || q > 3) return -1;
    ensure_helm(o, q, dt);
    size_t n = o->NT;
    const ora_mat *A = &o->helm[q];
    double *x = o->f[TB_F_U + q], *b = o->f[TB_F_RHS_U + q];
    double *r = dalloc(n), *rh = dalloc(n), *p = dalloc(n), *v = dalloc(n), *s = dalloc(n), *t = dalloc(n);
    ora_mat_apply(A, x, r);
    for (size_t i = 0; i < n; i++) { r[i] = b[i] - r[i]; rh[i] = r[i]; }
    double bn = sqrt(dot(b, b, n));
    if (bn == 0.0) bn = 1.0;
    double rn = sqrt(dot(r, r, n));
    double res0 = rn / bn;
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0, conv = (rn <= rtol * bn);
    while (!conv && it < maxit) {
        double rho1 = dot(rh, r, n);
        if (rho1 == 0.0) break;
        double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
        for (size_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        ora_mat_apply(A, p, v);
        alpha = rho / dot(rh, v, n);
        for (size_t i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        it++;
        double sn = sqrt(dot(s, s, n));
        if (sn <= rtol * bn) {
            for (size_t i = 0; i < n; i++) x[i] += alpha * p[i];
            rn = sn; conv = 1; break;
        }
        ora_mat_apply(A, s, t);
        double tt = dot(t, t, n);
        omega = (tt > 0.0) ? dot(t, s, n) / tt : 0.0;
        for (size_t i = 0; i < n; i++) { x[i] += alpha * p[i] + omega * s[i]; r[i] = s[i] - omega * t[i]; }
        rn = sqrt(dot(r, r, n));
        conv = (rn <= rtol * bn);
        if (omega == 0.0) break;
    }
    /* true residual */
    ora_mat_apply(A, x, r);
    for (size_t i = 0; i < n; i++) r[i] = b[i] - r[i];
    double tr = sqrt(dot(r, r, n)) / bn;
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iterations = it; st->converged = (tr <= rtol * 10.0); st->res0 = res0; st->res = tr;
    }
    free(r); free(rh); free(p); free(v); free(s); free(t);
    return (st && !st->converged) ? 1 : 0;
}

/* ---------------------------------------------------------------------- */
/* Extract_total_kinetic_energy (Extract.c:803-884) and                     */
/* Conc_compute_total_suspended_mass (Conc.c:1471-1512): serial k, j, i sums */
/* over FLUID cells of the concentration grid, as the reference loops them. */

int ora_diagnostics(ora_ctx *o, double *ke, double *mass) {
    cell_center(o);
    const double *c = o->f[TB_F_C];
    double e = 0.0, m = 0.0;
    for (int k = 0; k < o->NZ; k++) {
        double dz = o->dz_c[k];
        for (int j = 0; j < o->NY; j++) {
            double dy = o->dy_c[j];
            for (int i = 0; i < o->NX; i++) {
                if (ora_status(o, TB_Q_C, i, j, k) != TB_FLUID) continue;
                size_t id = ora_idx(o, i, j, k);
                double dx = o->dx_c[i];
                double dV = dx * dy * dz;
                double u_ = o->ubc[id], v_ = o->vbc[id], w_ = o->wbc[id];
                e += (u_ * u_ + v_ * v_ + w_ * w_) * dV;
                double dVc = dz * dy * o->dx_c[i];
                m += c[id] * dVc;
            }
        }
    }
    if (ke) *ke = 0.5 * e;
    if (mass) *mass = m;
    return 0;
}

/* Grid_set_interface_y_index (Grid.c:1972-2009) + Conc_bottom_sed_rate      */
/* (Conc.c:2158-2229) + Conc_compute_sedimentation_rate (Conc.c:2098-2157)  */
double ora_sedimentation_rate(ora_ctx *o, int particle, double v_settl, double *bottom) {
    double tot = 0.0;
    double us = fabs(v_settl);
    for (int k = 0; k < o->NZ; k++) {
        double dz = o->dz_c[k];
        for (int i = 0; i < o->NX; i++) {
            int jf = -1;
            for (int j = 0; j < o->NY; j++) {
                int s = ora_status(o, TB_Q_C, i, j, k);
                if (s == TB_IMMERSED) { jf = j; break; }
                if (j == 0 && s == TB_FLUID) { jf = j; break; }
            }
            double dx = o->dx_c[i];
            double r = 0.0;
            if (particle && jf >= 0) r = us * o->f[TB_F_C][ora_idx(o, i, jf, k)] * dx * dz;
            if (bottom) bottom[(size_t)k * o->NX + i] = r;
            if (particle && jf >= 0) tot += r;
        }
    }
    return tot;
}

/* Extract_Stokes_dissipation_rate (Extract.c:886-950) */
double ora_stokes_dissipation(ora_ctx *o, int particle, double v_settl, double conc_alpha) {
    if (!particle) return 0.0;
    const double *c = o->f[TB_F_C];
    double r = 0.0;
    for (int k = 0; k < o->NZ; k++) {
        double dz = o->dz_c[k];
        for (int j = 0; j < o->NY; j++) {
            double dy = o->dy_c[j];
            for (int i = 0; i < o->NX; i++) {
                if (ora_status(o, TB_Q_C, i, j, k) != TB_FLUID) continue;
                double dx = o->dx_c[i];
                double dV = dx * dy * dz;
                r += c[ora_idx(o, i, j, k)] * dV;
            }
        }
    }
    r *= fabs(v_settl) * conc_alpha;
    return r;
}

/* ---------------------------------------------------------------------- */
/* Pressure_compute_real_pressure (Pressure.c:590-685): the right-hand side  */
/* -div(grad p) from the stored pressure gradients, 0 on BOUNDARY cells;    */
/* the caller then solves the Poisson system into phi (Pressure_solve).     */

int ora_real_pressure_rhs(ora_ctx *o) {
    const double *dpx = o->f[TB_F_DPDX], *dpy = o->f[TB_F_DPDY], *dpz = o->f[TB_F_DPDZ];
    double *rhs = o->f[TB_F_PRHS];
    double a_dEta = 1.0 / o->dEta, a_dXi = 1.0 / o->dXi, a_dPhi = 1.0 / o->dPhi;
    size_t pl = (size_t)o->NX * o->NY;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++) {
        double C_z = o->mzc[k] * a_dPhi;
        for (int j = 0; j < o->NY; j++) {
            double C_y = o->myc[j] * a_dXi;
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                double r = 0.0;
                if (o->st[TB_Q_C][id] != TB_BOUNDARY) {
                    double C_x = o->mxc[i] * a_dEta;
                    double ddx = (dpx[id + 1] - dpx[id]) * C_x;
                    double ddy = (dpy[id + o->NX] - dpy[id]) * C_y;
                    double ddz = (dpz[id + pl] - dpz[id]) * C_z;
                    r = -(ddx + ddy + ddz);
                }
                rhs[id] = r;
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Poisson: Pressure_set_RHS (Pressure.c:171-276)                           */

int ora_poisson_rhs(ora_ctx *o, double dt) {
    const double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    double *rhs = o->f[TB_F_PRHS];
    double a_dEta = 1.0 / o->dEta, a_dXi = 1.0 / o->dXi, a_dPhi = 1.0 / o->dPhi;
    double a_dt = 1.0 / dt;
    size_t pl = (size_t)o->NX * o->NY;
    /* open x faces: u there from its plane (Pressure.c:242-249) */
    const double *uin = ora_in_plane(o, TB_Q_U), *uout = ora_out_plane(o, TB_Q_U);
    int NI = o->NX - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++) {
        double C_z = o->mzc[k] * a_dPhi;
        for (int j = 0; j < o->NY; j++) {
            double C_y = o->myc[j] * a_dXi;
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                double r = 0.0;
                if (o->st[TB_Q_C][id] != TB_BOUNDARY) {
                    double C_x = o->mxc[i] * a_dEta;
                    double dudx = (u[id + 1] - u[id]) * C_x;
                    if (uin && i == 0) dudx = (u[id + 1] - uin[ora_pidx(o, j, k)]) * C_x;
                    if (uout && i == NI - 1) dudx = (uout[ora_pidx(o, j, k)] - u[id]) * C_x;
                    double dvdy = (v[id + o->NX] - v[id]) * C_y;
                    double dwdz = (w[id + pl] - w[id]) * C_z;
                    r = -(dudx + dvdy + dwdz) * a_dt;
                }
                rhs[id] = r;
            }
        }
    }
    return 0;
}

/* Pressure_form_LHS_matrix (Pressure.c:691-936) */
static void build_pois(ora_ctx *o) {
    ora_mat *m = &o->pois;
    mat_alloc(m, o->NT);
    double Cx = -1.0 / (o->dEta * o->dEta), Cy = -1.0 / (o->dXi * o->dXi), Cz = -1.0 / (o->dPhi * o->dPhi);
    double ap_solid = -Cx * o->mxc[0] - Cy * o->myc[0] - Cz * o->mzc[0];
    int NI = o->NX - 1, NJ = o->NY - 1, NK = o->NZ - 1;
    size_t pl = (size_t)o->NX * o->NY;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                int64_t cols[9];
                double vals[9];
                int nnz = 0;
                if (o->st[TB_Q_C][id] != TB_BOUNDARY) {
                    double mpx = o->mxc[i], mpy = o->myc[j], mpz = o->mzc[k];
                    double mex = o->mxu[i + 1], mwx = o->mxu[i], mny = o->myv[j + 1], msy = o->myv[j];
                    double mfz = o->mzw[k + 1], mbz = o->mzw[k];
                    double ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
                    double aw = Cx * mpx * mwx, ae = Cx * mpx * mex, an = Cy * mpy * mny;
                    double as = Cy * mpy * msy, ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
#define ADD(v, c) do { vals[nnz] = (v); cols[nnz] = (int64_t)(c); nnz++; } while (0)
                    if (k > 0) ADD(ab, id - pl); else ap += ab;
                    if (j > 0) ADD(as, id - o->NX); else ap += as;
                    if (i > 0) ADD(aw, id - 1); else ap += aw;
                    if (i < NI - 1) ADD(ae, id + 1);
                    else if (!((i == NI - 1) && (j == NJ - 1) && (k == NK - 1))) ap += ae;
                    if (j < NJ - 1) ADD(an, id + o->NX); else ap += an;
                    if (k < NK - 1) ADD(af, id + pl); else ap += af;
                    ADD(ap, id);
                } else {
                    ADD(ap_solid, id);
                }
#undef ADD
                row_set(m, id, nnz, cols, vals);
            }
    o->pois_built = 1;
}

int ora_poisson_apply(ora_ctx *o, const double *x, double *y) {
    if (!o->pois_built) build_pois(o);
    ora_mat_apply(&o->pois, x, y);
    return 0;
}

/* interior cell <-> compact array */
static void gather_cells(const ora_ctx *o, const double *pad, double *cmp) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->ncz; k++)
        for (int j = 0; j < o->ncy; j++)
            for (int i = 0; i < o->ncx; i++)
                cmp[((size_t)k * o->ncy + j) * o->ncx + i] = pad[ora_idx(o, i, j, k)];
}

static void scatter_cells(const ora_ctx *o, const double *cmp, double *pad) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->ncz; k++)
        for (int j = 0; j < o->ncy; j++)
            for (int i = 0; i < o->ncx; i++)
                pad[ora_idx(o, i, j, k)] = cmp[((size_t)k * o->ncy + j) * o->ncx + i];
}

/* Pressure_solve replacement: CG on the Neumann operator (the reference
 * operator without its corner pin), preconditioned by a symmetric GMG
 * V-cycle, in the mean-free subspace; the result is shifted so that the
 * pinned corner cell holds 0, which makes it the solution of the reference's
 * pinned system (Pressure.c:864-868) for a compatible right-hand side.
 * Convergence is judged by the true residual of the pinned operator. */
int ora_poisson_solve(ora_ctx *o, double rtol, int maxit, tb_solve_stats *st) {
    if (!o->pois_built) build_pois(o);
    if (o->nlev == 0) ora_gmg_build(o);
    size_t n = o->NT;
    size_t nc = (size_t)o->ncx * o->ncy * o->ncz;
    double *x = o->f[TB_F_PHI], *b = o->f[TB_F_PRHS];
    double *r = dalloc(n), *z = dalloc(n), *p = dalloc(n), *qv = dalloc(n);
    double *rc = dalloc(nc), *zc = dalloc(nc);
    size_t corner = ora_idx(o, o->ncx - 1, o->ncy - 1, o->ncz - 1);
    /* pin coefficient: the east coefficient the corner row keeps */
    double Cx = -1.0 / (o->dEta * o->dEta);
    double pin = -(Cx * o->mxc[o->ncx - 1] * o->mxu[o->ncx]);
    double bn = sqrt(dot(b, b, n));
    int it = 0, vcy = 0;
    double res0 = 0.0, res = 0.0;
    if (bn == 0.0) {
        memset(x, 0, sizeof(double) * n);
        if (st) { memset(st, 0, sizeof(*st)); st->converged = 1; }
        goto done;
    }
    /* boundary (half-cell) entries of phi are 0 (ap_solid*phi = 0 rows) */
    for (size_t id = 0; id < n; id++) if (o->st[TB_Q_C][id] == TB_BOUNDARY) x[id] = 0.0;
    {
        double xc0 = x[corner];
        for (size_t id = 0; id < n; id++) if (o->st[TB_Q_C][id] != TB_BOUNDARY) x[id] -= xc0;
    }
    ora_mat_apply(&o->pois, x, r);
    for (size_t id = 0; id < n; id++) r[id] = b[id] - r[id];
    res0 = sqrt(dot(r, r, n)) / bn;
    res = res0;
    while (res > rtol && it < maxit) {
        /* residual of the Neumann operator: A_N x = A_p x - pin*x_c e_c, and
         * x_c = 0 here, so the pinned residual is also the Neumann one */
        gather_cells(o, r, rc);
        double mean = 0.0;
        for (size_t a = 0; a < nc; a++) mean += rc[a];
        mean /= (double)nc;
        for (size_t a = 0; a < nc; a++) rc[a] -= mean;
        scatter_cells(o, rc, r);
        ora_vcycle(o, rc, zc); vcy++;
        memset(z, 0, sizeof(double) * n);
        scatter_cells(o, zc, z);
        memcpy(p, z, sizeof(double) * n);
        double rz = dot(r, z, n);
        int inner = 0;
        while (it < maxit) {
            ora_mat_apply(&o->pois, p, qv);
            qv[corner] -= pin * p[corner];        /* Neumann operator */
            double alpha = rz / dot(p, qv, n);
            for (size_t id = 0; id < n; id++) { x[id] += alpha * p[id]; r[id] -= alpha * qv[id]; }
            it++; inner++;
            double rr = sqrt(dot(r, r, n)) / bn;
            if (rr <= 0.5 * rtol || inner >= 200) break;
            gather_cells(o, r, rc);
            ora_vcycle(o, rc, zc); vcy++;
            scatter_cells(o, zc, z);
            double rz1 = dot(r, z, n);
            double beta = rz1 / rz;
            rz = rz1;
            for (size_t id = 0; id < n; id++) p[id] = z[id] + beta * p[id];
        }
        double xc0 = x[corner];
        for (size_t id = 0; id < n; id++) if (o->st[TB_Q_C][id] != TB_BOUNDARY) x[id] -= xc0;
        ora_mat_apply(&o->pois, x, r);
        for (size_t id = 0; id < n; id++) r[id] = b[id] - r[id];
        res = sqrt(dot(r, r, n)) / bn;
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iterations = it; st->vcycles = vcy; st->converged = (res <= rtol);
        st->res0 = res0; st->res = res;
    }
done:
    free(r); free(z); free(p); free(qv); free(rc); free(zc);
    return (st && !st->converged) ? 1 : 0;
}

/* Pressure_project_velocity (Pressure.c:280-439) */
int ora_project(ora_ctx *o, double dt) {
    double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    double *dpx = o->f[TB_F_DPDX], *dpy = o->f[TB_F_DPDY], *dpz = o->f[TB_F_DPDZ];
    const double *phi = o->f[TB_F_PHI], *b = o->f[TB_F_PRHS];
    double coef = dt / o->cfg.re;
    double a_dEta = 1.0 / o->dEta, a_dXi = 1.0 / o->dXi, a_dPhi = 1.0 / o->dPhi;
    size_t pl = (size_t)o->NX * o->NY;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                if (o->st[TB_Q_U][id] != TB_BOUNDARY) {
                    double metric_ex = o->mxu[i];
                    double d = (phi[id] - phi[id - 1]) * a_dEta * metric_ex;
                    u[id] -= dt * d;
                    double lap = -(b[id] - b[id - 1]) * metric_ex * a_dEta;
                    dpx[id] += d - coef * lap;
                }
                if (o->st[TB_Q_V][id] != TB_BOUNDARY) {
                    double C_y = a_dXi * o->myv[j];
                    double d = (phi[id] - phi[id - o->NX]) * C_y;
                    v[id] -= dt * d;
                    double lap = -(b[id] - b[id - o->NX]) * C_y;
                    dpy[id] += d - coef * lap;
                }
                if (o->st[TB_Q_W][id] != TB_BOUNDARY) {
                    double C_z = o->mzw[k] * a_dPhi;
                    double d = (phi[id] - phi[id - pl]) * C_z;
                    w[id] -= dt * d;
                    double lap = -(b[id] - b[id - pl]) * C_z;
                    dpz[id] += d - coef * lap;
                }
            }
    return 0;
}

/* Pressure_compute_velocity_divergence (Pressure.c:462-586) */
double ora_div_max(ora_ctx *o, int32_t *ijk) {
    const double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    double *div = o->f[TB_F_DIV];
    double a_dEta = 1.0 / o->dEta, a_dXi = 1.0 / o->dXi, a_dPhi = 1.0 / o->dPhi;
    size_t pl = (size_t)o->NX * o->NY;
    double div_max = 0.0;
    int im = -1, jm = -1, km = -1;
    /* open x faces: u there from its plane (Pressure.c:541-548) */
    const double *uin = ora_in_plane(o, TB_Q_U), *uout = ora_out_plane(o, TB_Q_U);
    int NI = o->NX - 1;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                double dc = 0.0;
                if (o->st[TB_Q_C][id] != TB_BOUNDARY) {
                    double dudx = (u[id + 1] - u[id]) * a_dEta * o->mxc[i];
                    if (uin && i == 0) dudx = (u[id + 1] - uin[ora_pidx(o, j, k)]) * a_dEta * o->mxc[i];
                    if (uout && i == NI - 1) dudx = (uout[ora_pidx(o, j, k)] - u[id]) * a_dEta * o->mxc[i];
                    double dvdy = (v[id + o->NX] - v[id]) * a_dXi * o->myc[j];
                    double dwdz = (w[id + pl] - w[id]) * a_dPhi * o->mzc[k];
                    dc = fabs(dudx + dvdy + dwdz);
                }
                div[id] = dc;
                if (dc >= div_max) { div_max = dc; im = i; jm = j; km = k; }
            }
    if (ijk) { ijk[0] = im; ijk[1] = jm; ijk[2] = km; }
    return div_max;
}

int ora_store_old(ora_ctx *o) {
    for (int q = 0; q < 3; q++) memcpy(o->old[q], o->f[TB_F_U + q], sizeof(double) * o->NT);
    for (int n = 0; n < o->nconc; n++) memcpy(o->cs[n].old, o->cs[n].c, sizeof(double) * o->NT);
    ora_open_store_old(o);
    return 0;
}

/* CFL time step.  The reference's parameters are cfl_method, alpha and
 * default_dt (Input.c:421-423); its rule lived in the unpublished driver.
 * Restated rule (the same as the device's): over FLUID cells of the
 * concentration grid, with the cell-centre velocities of
 * Velocity_cell_center (Velocity.c:1476-1602),
 * rate = |u_c|/dx + |v_c|/dy + |w_c|/dz; dt = min(dt_max, alpha / max rate). */
double ora_cfl_dt(ora_ctx *o, double alpha, double dt_max, double *rate) {
    cell_center(o);
    double m = 0.0;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t id = ora_idx(o, i, j, k);
                if (o->st[TB_Q_C][id] != TB_FLUID) continue;
                double r = fabs(o->ubc[id]) / o->dx_c[i] + fabs(o->vbc[id]) / o->dy_c[j] +
                           fabs(o->wbc[id]) / o->dz_c[k];
                if (r != r || r > m) m = r;
            }
    if (rate) *rate = m;
    return (m > 0.0 && alpha / m < dt_max) ? alpha / m : dt_max;
}

/* VecAXPBY(y, a, b, x): y = a*x + b*y  (Velocity.c:2047-2055, Conc.c:1007-1017) */
int ora_rk_average(ora_ctx *o, double old_frac, double new_frac) {
    for (int q = 0; q < 3 + o->nconc; q++) {
        double *y = (q < 3) ? o->f[TB_F_U + q] : o->cs[q - 3].c;
        const double *x = (q < 3) ? o->old[q] : o->cs[q - 3].old;
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < o->NT; i++) y[i] = old_frac * x[i] + new_frac * y[i];
    }
    ora_open_rk_average(o, old_frac, new_frac);
    return 0;
}

/* SSP-RK2 step (TVD_RK2, definitions.h:148-149) in the reconstructed order:
 * explicit terms of every field from the stage-start state, then the four
 * Helmholtz solves, then Poisson + projection; after two stages the
 * 0.5/0.5 RK average of u, v, w, c. */
int ora_step(ora_ctx *o, double dt, tb_step_stats *st) {
    if (st) memset(st, 0, sizeof(*st));
    ora_store_old(o);
    int sel = o->conc_sel;
    for (int stage = 0; stage < 2; stage++) {
        /* the convective outflow condition from the stage-start fields */
        if (o->obc.outflow) ora_outflow_convective(o, dt);
        ora_total_concentration(o);
        for (int q = 0; q < 3; q++) ora_eno_conv(o, q);
        for (int n = 0; n < o->nconc; n++) {
            ora_select_conc(o, n);
            ora_eno_conv(o, TB_Q_C);
            ora_set_rhs(o, TB_Q_C, dt);
        }
        ora_select_conc(o, sel);
        for (int q = 0; q < 3; q++) ora_set_rhs(o, q, dt);
        for (int q = 0; q < 3; q++)
            ora_helmholtz_solve(o, q, dt, o->cfg.helm_rtol, o->cfg.helm_maxit,
                                st ? &st->helm[stage][q] : NULL);
        for (int n = 0; n < o->nconc; n++) {
            tb_solve_stats cst;
            ora_select_conc(o, n);
            ora_helmholtz_solve(o, TB_Q_C, dt, o->cfg.helm_rtol, o->cfg.helm_maxit, &cst);
            if (st && n == 0) st->helm[stage][TB_Q_C] = cst;
        }
        ora_select_conc(o, sel);
        ora_poisson_rhs(o, dt);
        ora_poisson_solve(o, o->cfg.pois_rtol, o->cfg.pois_maxit, st ? &st->pois[stage] : NULL);
        ora_project(o, dt);
    }
    ora_rk_average(o, 0.5, 0.5);
    double dm = ora_div_max(o, NULL);
    if (st) st->div_max = dm;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* concentration fields                                                     */

/* Conc_settling_speed_function (Conc.c:815-836): V_s0 (1 - c phi_max)^n with
 * n = 4 (the Richardson-Zaki form the reference returns), the power taken as
 * two squarings */
double ora_hindered_speed(double conc, double phi_max, double v_s0) {
    double t = 1.0 - conc * phi_max;
    double t2 = t * t;
    return v_s0 * (t2 * t2);
}

int ora_select_conc(ora_ctx *o, int n) {
    if (n < 0 || n >= o->nconc) return -1;
    const ora_conc *s = &o->cs[n];
    o->f[TB_F_C] = s->c; o->f[TB_F_CONV_C] = s->conv; o->f[TB_F_RHS_C] = s->rhs; o->old[3] = s->old;
    if (o->cfg.pe != s->p.pe || o->cfg.conc_type != s->p.type || o->cfg.v_settle != s->p.v_settle)
        o->helm_dt[TB_Q_C] = 0.0;   /* the concentration rows depend on these */
    o->cfg.pe = s->p.pe;
    o->cfg.conc_type = s->p.type;
    o->cfg.v_settle = s->p.v_settle;
    o->conc_sel = n;
    return 0;
}

int ora_set_conc_fields(ora_ctx *o, int nconc, const tb_conc_field *fields, int hindered) {
    if (nconc < 1 || nconc > TB_NCONC_MAX || !fields) return -1;
    ora_select_conc(o, 0);
    for (int n = 1; n < TB_NCONC_MAX; n++) {
        ora_conc *s = &o->cs[n];
        if (n < nconc && !s->c) {
            s->c = dalloc(o->NT); s->conv = dalloc(o->NT); s->rhs = dalloc(o->NT); s->old = dalloc(o->NT);
        } else if (n >= nconc && s->c) {
            free(s->c); free(s->conv); free(s->rhs); free(s->old); free(s->dep);
            s->c = s->conv = s->rhs = s->old = s->dep = NULL;
        }
    }
    if (nconc > 1 && !o->ctot) o->ctot = dalloc(o->NT);
    if (nconc == 1) { free(o->ctot); o->ctot = NULL; }
    for (int n = 0; n < nconc; n++) o->cs[n].p = fields[n];
    o->nconc = nconc;
    o->hindered = hindered != 0;
    return ora_select_conc(o, 0);
}

/* Conc_compute_total_concentration (Conc.c:1291-1314): VecCopy + VecScale by
 * alpha_0, then VecAXPY alpha_i c_i for i = 1.. */
int ora_total_concentration(ora_ctx *o) {
    if (o->nconc == 1) return 0;
    for (size_t id = 0; id < o->NT; id++) {
        double t = o->cs[0].c[id];
        t = t * o->cs[0].p.alpha;
        for (int n = 1; n < o->nconc; n++) t = t + o->cs[n].p.alpha * o->cs[n].c[id];
        o->ctot[id] = t;
    }
    return 0;
}
