/*
 * ora_open.c -- oracle restatement of the open x boundaries: the inflow
 * profiles (Inflow.c:13-260), the convective outflow condition with the
 * global mass correction (Outflow.c:12-304) and the outflow planes' RK
 * bookkeeping (Velocity.c:2058-2120, Conc.c:1223-1289).  The operators that
 * read the planes (ghosts, transposes, cell centres, right-hand sides,
 * divergence, Helmholtz rows) take them in ora_flow.c / ora_diag.c.
 * TEST INFRASTRUCTURE ONLY (see tb_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

static size_t plane_n(const ora_ctx *o) { return (size_t)o->NY * (size_t)o->NZ; }

const double *ora_in_plane(const ora_ctx *o, int q) {
    if (!o->obc.inflow || !o->u_in) return NULL;
    if (q == TB_Q_U) return o->u_in;
    if (q == TB_Q_C) return o->cs[o->conc_sel].cin;
    return NULL;
}

const double *ora_out_plane(const ora_ctx *o, int q) {
    if (!o->obc.outflow || !o->out[0]) return NULL;
    if (q < 3) return o->out[q];
    return o->cs[o->conc_sel].cout;
}

void ora_open_free(ora_ctx *o) {
    free(o->u_in);
    o->u_in = NULL;
    for (int q = 0; q < 3; q++) {
        free(o->out[q]); free(o->out_old[q]);
        o->out[q] = o->out_old[q] = NULL;
    }
    for (int n = 0; n < TB_NCONC_MAX; n++) {
        free(o->cs[n].cin); free(o->cs[n].cout); free(o->cs[n].cout_old);
        o->cs[n].cin = o->cs[n].cout = o->cs[n].cout_old = NULL;
    }
}

int ora_set_open_boundaries(ora_ctx *o, const tb_open_bc *bc) {
    if (!o || !bc || (bc->inflow != 0 && bc->inflow != 1) || (bc->outflow != 0 && bc->outflow != 1) ||
        !isfinite(bc->u_conv) || (bc->transient_inflow && !(bc->time_c > 0.0)))
        return -1;
    ora_open_free(o);
    const size_t n = plane_n(o);
    o->u_in = (double *)calloc(n, sizeof(double));
    for (int q = 0; q < 3; q++) {
        o->out[q] = (double *)calloc(n, sizeof(double));
        o->out_old[q] = (double *)calloc(n, sizeof(double));
    }
    for (int s = 0; s < TB_NCONC_MAX; s++) {
        o->cs[s].cin = (double *)calloc(n, sizeof(double));
        o->cs[s].cout = (double *)calloc(n, sizeof(double));
        o->cs[s].cout_old = (double *)calloc(n, sizeof(double));
    }
    o->obc = *bc;
    o->influx = 0.0;
    for (int q = 0; q < 4; q++) o->helm_dt[q] = 0.0;   /* the rows at the open faces change */
    return 0;
}

/* sum of plane * dy dz over the FLUID u nodes of i = 1 (Inflow.c:120-139) */
static double u_influx(const ora_ctx *o, const double *plane) {
    double flux = 0.0;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            if (ora_status(o, TB_Q_U, 1, j, k) == TB_FLUID) {
                double dy = o->yv[j + 1] - o->yv[j];
                double dz = o->zw[k + 1] - o->zw[k];
                double dA = dy * dz;
                flux += plane[ora_pidx(o, j, k)] * dA;
            }
    return flux;
}

int ora_inflow_profile(ora_ctx *o, double time, double *influx) {
    if (!o || !o->u_in) return -1;
    double tf = o->obc.transient_inflow ? (1.0 - exp(-(time + 0.01) / (o->obc.time_c))) : 1.0;
    /* Inflow_u_velocity_profile: u_inflow = 1.0 (the final assignment,
     * Inflow.c:99), times the time factor, on FLUID u nodes of i = 1 */
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            if (ora_status(o, TB_Q_U, 1, j, k) == TB_FLUID) {
                double v = 1.0;
                v *= tf;
                o->u_in[ora_pidx(o, j, k)] = v;
            }
    o->influx = u_influx(o, o->u_in);
    /* Inflow_conc_profile for every field: c_inflow = 1.0 (Inflow.c:233)
     * times the time factor on the non-SOLID cells of i = 0 */
    for (int s = 0; s < o->nconc; s++)
        for (int k = 0; k < o->NZ; k++)
            for (int j = 0; j < o->NY; j++)
                if (ora_status(o, TB_Q_C, 0, j, k) != TB_SOLID) {
                    double v = 1.0;
                    v *= tf;
                    o->cs[s].cin[ora_pidx(o, j, k)] = v;
                }
    if (influx) *influx = o->influx;
    return 0;
}

static double *plane_of(ora_ctx *o, int which, int q) {
    if (!o->u_in || q < 0 || q > 3) return NULL;
    if (which == TB_PLANE_INFLOW) {
        if (q == TB_Q_U) return o->u_in;
        if (q == TB_Q_C) return o->cs[o->conc_sel].cin;
        return NULL;
    }
    if (which == TB_PLANE_OUTFLOW) return q < 3 ? o->out[q] : o->cs[o->conc_sel].cout;
    return NULL;
}

int ora_set_boundary_plane(ora_ctx *o, int which, int q, const double *plane) {
    double *p = o ? plane_of(o, which, q) : NULL;
    if (!p || !plane) return -1;
    memcpy(p, plane, sizeof(double) * plane_n(o));
    if (which == TB_PLANE_INFLOW && q == TB_Q_U) o->influx = u_influx(o, p);
    return 0;
}

int ora_get_boundary_plane(ora_ctx *o, int which, int q, double *plane) {
    double *p = o ? plane_of(o, which, q) : NULL;
    if (!p || !plane) return -1;
    memcpy(plane, p, sizeof(double) * plane_n(o));
    return 0;
}

double ora_get_influx(ora_ctx *o) { return o ? o->influx : 0.0; }

/* Outflow_vel_impose_convective_boundary / Outflow_conc_impose_convective_boundary
 * for one plane: out = (1 - coef) out + coef q[NI-1] on rows whose node at
 * i = NI-1 is FLUID or IMMERSED (Outflow.c:44-87, 120-156) */
static void convect(const ora_ctx *o, int q, const double *data, double *out, double coef) {
    int last = o->NX - 2;   /* NI - 1 with NI = grid->NI = NX - 1 */
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++) {
            int s = ora_status(o, q, last, j, k);
            if (s == TB_FLUID || s == TB_IMMERSED) {
                size_t p = ora_pidx(o, j, k);
                out[p] = (1.0 - coef) * out[p] + coef * data[ora_idx(o, last, j, k)];
            }
        }
}

/* Outflow_update_u_velocity_to_conserve_mass (Outflow.c:163-240): outflux
 * and area over the open rows, then the uniform correction */
static void conserve_mass(ora_ctx *o) {
    int last = o->NX - 2;
    double outflux = 0.0, area = 0.0;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++) {
            int s = ora_status(o, TB_Q_U, last, j, k);
            if (s == TB_FLUID || s == TB_IMMERSED) {
                double dy = o->yv[j + 1] - o->yv[j];
                double dz = o->zw[k + 1] - o->zw[k];
                double dA = dy * dz;
                area += dA;
                outflux += dA * o->out[0][ora_pidx(o, j, k)];
            }
        }
    if (!(area > 0.0)) return;   /* no open row: nothing to correct */
    double corr = (o->influx - outflux) / area;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++) {
            int s = ora_status(o, TB_Q_U, last, j, k);
            if (s == TB_FLUID || s == TB_IMMERSED) o->out[0][ora_pidx(o, j, k)] += corr;
        }
}

int ora_outflow_convective(ora_ctx *o, double dt) {
    if (!o || !o->obc.outflow || !o->out[0] || !(dt > 0.0)) return -1;
    int NI = o->NX - 1;
    double dx = o->xu[NI] - o->xu[NI - 1];
    double coef = dt * o->obc.u_conv / dx;
    convect(o, TB_Q_U, o->f[TB_F_U], o->out[0], coef);
    conserve_mass(o);
    convect(o, TB_Q_V, o->f[TB_F_V], o->out[1], coef);
    convect(o, TB_Q_W, o->f[TB_F_W], o->out[2], coef);
    for (int s = 0; s < o->nconc; s++) convect(o, TB_Q_C, o->cs[s].c, o->cs[s].cout, coef);
    return 0;
}

/* Velocity_store_old_outflow / Conc_store_old_outflow */
void ora_open_store_old(ora_ctx *o) {
    if (!o->obc.outflow || !o->out[0]) return;
    size_t n = plane_n(o);
    for (int q = 0; q < 3; q++) memcpy(o->out_old[q], o->out[q], sizeof(double) * n);
    for (int s = 0; s < o->nconc; s++) memcpy(o->cs[s].cout_old, o->cs[s].cout, sizeof(double) * n);
}

/* Velocity_rk_average_outflow / Conc_rk_average_outflow */
void ora_open_rk_average(ora_ctx *o, double old_frac, double new_frac) {
    if (!o->obc.outflow || !o->out[0]) return;
    size_t n = plane_n(o);
    for (int q = 0; q < 4 + o->nconc - 1; q++) {
        double *a = q < 3 ? o->out[q] : o->cs[q - 3].cout;
        const double *b = q < 3 ? o->out_old[q] : o->cs[q - 3].cout_old;
        for (size_t p = 0; p < n; p++) a[p] = old_frac * b[p] + new_frac * a[p];
    }
}
