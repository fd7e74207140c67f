/*
 * ora_diag.c -- oracle restatement of the flow diagnostics of the
 * reference's output stage (Output_flow_properties, Output.c:2062-2216):
 * viscous dissipation, kinetic/potential energy conversion, potential
 * energies, bottom wall shear (linear and quadratic), span-averaged height,
 * front location and deposited height.  Sums run in the reference's serial
 * loop order.  TEST INFRASTRUCTURE ONLY (see tb_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

/* Grid_set_interface_y_index (Grid.c:1972-2009): the first IMMERSED cell of
 * column (i, k) from the bottom, or row 0 when it is FLUID, else NOT_FOUND */
int ora_interface_y_index(const ora_ctx *o, int i, int k) {
    for (int j = 0; j < o->NY; j++) {
        int s = ora_status(o, TB_Q_C, i, j, k);
        if (s == TB_IMMERSED) return j;
        if (j == 0 && s == TB_FLUID) return j;
    }
    return ORA_NOT_FOUND;
}

/* Boundary.h selects each wall at compile time; Extract.c:487-522 turns the
 * choice into the ghost factor: no-slip -1, free-slip +1 */
static double ghost_factor(const ora_ctx *o, int wall) { return o->cfg.wall_noslip[wall] ? -1.0 : 1.0; }

/* Extract_viscous_dissipation_rate (Extract.c:453-725): (1/Re) sum over
 * FLUID cells of u lap(u) dV_u + v lap(v) dV_v + w lap(w) dV_w, the
 * Laplacians with the wall ghosts u_ghost = k u; an open x face gives the
 * neighbours there from its planes (Extract.c:577-590, 630-645) */
double ora_viscous_dissipation(ora_ctx *o) {
    const double *u = o->f[TB_F_U], *v = o->f[TB_F_V], *w = o->f[TB_F_W];
    const double k_left = ghost_factor(o, 0), k_right = ghost_factor(o, 1);
    const double k_bottom = ghost_factor(o, 2), k_top = ghost_factor(o, 3);
    const double k_back = ghost_factor(o, 4), k_front = ghost_factor(o, 5);
    const double Cx = 1.0 / (o->dEta * o->dEta), Cy = 1.0 / (o->dXi * o->dXi), Cz = 1.0 / (o->dPhi * o->dPhi);
    const int NI = o->NX - 1, NJ = o->NY - 1, NK = o->NZ - 1;
    const size_t pl = (size_t)o->NX * o->NY, rx = (size_t)o->NX;
    const double *mxu = o->mxu, *mxc = o->mxc, *myv = o->myv, *myc = o->myc, *mzw = o->mzw, *mzc = o->mzc;
    const double *uin = ora_in_plane(o, TB_Q_U);
    const double *uout = ora_out_plane(o, TB_Q_U), *vout = ora_out_plane(o, TB_Q_V),
                 *wout = ora_out_plane(o, TB_Q_W);
    double rate = 0.0;
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                if (ora_status(o, TB_Q_C, i, j, k) != TB_FLUID) continue;
                size_t id = ora_idx(o, i, j, k);
                double U_ae = Cx * mxu[i] * mxc[i], U_aw = Cx * mxu[i];
                double U_an = Cy * myc[j] * myv[j + 1], U_as = Cy * myc[j] * myv[j];
                double U_af = Cz * mzc[k] * mzw[k + 1], U_ab = Cz * mzc[k] * mzw[k];
                double V_ae = Cx * mxc[i] * mxu[i + 1], V_aw = Cx * mxc[i] * mxu[i];
                double V_an = Cy * myv[j] * myc[j], V_as = Cy * myv[j];
                double V_af = Cz * mzc[k] * mzw[k + 1], V_ab = Cz * mzc[k] * mzw[k];
                double W_ae = Cx * mxc[i] * mxu[i + 1], W_aw = Cx * mxc[i] * mxu[i];
                double W_an = Cy * myc[j] * myv[j + 1], W_as = Cy * myc[j] * myv[j];
                double W_af = Cz * mzw[k] * mzc[k], W_ab = Cz * mzw[k];
                double U_w, V_w, W_w, U_e, V_e, W_e, U_s, V_s, W_s, U_n, V_n, W_n, U_b, V_b, W_b, U_f, V_f, W_f;
                if (i == 0) {
                    U_aw *= mxc[i];
                    U_w = uin ? uin[ora_pidx(o, j, k)] : 0.0;
                    V_w = k_left * v[id]; W_w = k_left * w[id];
                } else {
                    U_aw *= mxc[i - 1];
                    U_w = u[id - 1]; V_w = v[id - 1]; W_w = w[id - 1];
                }
                if (j == 0) {
                    V_as *= myc[j];
                    U_s = k_bottom * u[id]; V_s = 0.0; W_s = k_bottom * w[id];
                } else {
                    V_as *= myc[j - 1];
                    U_s = u[id - rx]; V_s = v[id - rx]; W_s = w[id - rx];
                }
                if (k == 0) {
                    W_ab *= mzc[k];
                    U_b = k_back * u[id]; V_b = k_back * v[id]; W_b = 0.0;
                } else {
                    W_ab *= mzc[k - 1];
                    U_b = u[id - pl]; V_b = v[id - pl]; W_b = w[id - pl];
                }
                if (i == NI - 1) {
                    if (uout) {
                        size_t p = ora_pidx(o, j, k);
                        U_e = uout[p]; V_e = vout[p]; W_e = wout[p];
                    } else {
                        U_e = 0.0; V_e = k_right * v[id]; W_e = k_right * w[id];
                    }
                } else {
                    U_e = u[id + 1]; V_e = v[id + 1]; W_e = w[id + 1];
                }
                if (j == NJ - 1) {
                    U_n = k_top * u[id]; V_n = 0.0; W_n = k_top * w[id];
                } else {
                    U_n = u[id + rx]; V_n = v[id + rx]; W_n = w[id + rx];
                }
                if (k == NK - 1) {
                    U_f = k_front * u[id]; V_f = k_front * v[id]; W_f = 0.0;
                } else {
                    U_f = u[id + pl]; V_f = v[id + pl]; W_f = w[id + pl];
                }
                double u_ = u[id], v_ = v[id], w_ = w[id];
                double U_ap = -(U_ae + U_aw + U_ab + U_af + U_as + U_an);
                double V_ap = -(V_ae + V_aw + V_ab + V_af + V_as + V_an);
                double W_ap = -(W_ae + W_aw + W_ab + W_af + W_as + W_an);
                double lu = (U_ae * U_e + U_aw * U_w) + (U_an * U_n + U_as * U_s) + (U_af * U_f + U_ab * U_b) + U_ap * u_;
                double lv = (V_ae * V_e + V_aw * V_w) + (V_an * V_n + V_as * V_s) + (V_af * V_f + V_ab * V_b) + V_ap * v_;
                double lw = (W_ae * W_e + W_aw * W_w) + (W_an * W_n + W_as * W_s) + (W_af * W_f + W_ab * W_b) + W_ap * w_;
                double dV_u = (o->xc[i + 1] - o->xc[i]) * o->dy_c[j] * o->dz_c[k];
                double dV_v = (o->yc[j + 1] - o->yc[j]) * o->dx_c[i] * o->dz_c[k];
                double dV_w = (o->zc[k + 1] - o->zc[k]) * o->dx_c[i] * o->dy_c[j];
                rate += u_ * lu * dV_u + v_ * lv * dV_v + w_ * lw * dV_w;
            }
    rate *= 1.0 / o->cfg.re;
    return rate;
}

/* the cell-centre v of Velocity_cell_center (Velocity.c:1476-1602) */
static double v_center(const ora_ctx *o, int i, int j, int k) {
    size_t id = ora_idx(o, i, j, k);
    const double *v = o->f[TB_F_V];
    return ora_interp2(v[id], o->yv[j], v[id + o->NX], o->yv[j + 1], o->yc[j]);
}

static double u_center(const ora_ctx *o, int i, int j, int k) {
    size_t id = ora_idx(o, i, j, k);
    const double *u = o->f[TB_F_U];
    return ora_interp2(u[id], o->xu[i], u[id + 1], o->xu[i + 1], o->xc[i]);
}

static double w_center(const ora_ctx *o, int i, int j, int k) {
    size_t id = ora_idx(o, i, j, k);
    const double *w = o->f[TB_F_W];
    return ora_interp2(w[id], o->zw[k], w[id + (size_t)o->NX * o->NY], o->zw[k + 1], o->zc[k]);
}

/* Extract_kinetic_potential_conversion_rate (Extract.c:727-802): sum over
 * FLUID cells of -v_c c_total dV, c_total from Conc_compute_total_
 * concentration with several fields (recomputed here), else c[0] */
double ora_kinetic_potential_conversion(ora_ctx *o) {
    ora_total_concentration(o);
    const double *ct = (o->nconc > 1) ? o->ctot : o->cs[0].c;
    double rate = 0.0;
    for (int k = 0; k < o->NZ; k++) {
        double dz = o->dz_c[k];
        for (int j = 0; j < o->NY; j++) {
            double dy = o->dy_c[j];
            for (int i = 0; i < o->NX; i++) {
                if (ora_status(o, TB_Q_C, i, j, k) != TB_FLUID) continue;
                double dx = o->dx_c[i];
                double v_ = v_center(o, i, j, k);
                double c_t = ct[ora_idx(o, i, j, k)];
                rate += -v_ * c_t * (dx * dy * dz);
            }
        }
    }
    return rate;
}

/* the selected field's settling speed as the reference keeps it
 * (v_settl0, 0 for a non-particle field, Conc.c:169-185) */
static double field_vs(const ora_ctx *o) { return o->cfg.conc_type == TB_PARTICLE ? o->cfg.v_settle : 0.0; }

/* Conc_compute_potential_energies (Conc.c:1934-2058) of the selected field:
 * out = {Ep, Ep_passive, Stokes dissipation, conversion, Ep_passive_rate,
 * Ep_rate} */
int ora_potential_energies(ora_ctx *o, double *out) {
    const double *c = o->f[TB_F_C];
    const double y_ref = o->yv[0];
    const int particle = o->cfg.conc_type == TB_PARTICLE;
    const double alpha = o->cs[o->conc_sel].p.alpha;
    const double *dep = o->cs[o->conc_sel].dep;
    double ep_passive = 0.0;
    if (particle && dep) {
        for (int k = 0; k < o->NZ; k++) {
            double dz = o->dz_c[k];
            for (int i = 0; i < o->NX; i++) {
                double dx = o->dx_c[i];
                int j = ora_interface_y_index(o, i, k);
                if (j >= 0) ep_passive += dep[(size_t)k * o->NX + i] * dx * dz * (o->yc[j] - y_ref);
            }
        }
    }
    double ep = 0.0, stokes = 0.0, conv = 0.0;
    for (int k = 0; k < o->NZ; k++) {
        double dz = o->dz_c[k];
        for (int j = 0; j < o->NY; j++) {
            double dy = o->dy_c[j];
            for (int i = 0; i < o->NX; i++) {
                if (ora_status(o, TB_Q_C, i, j, k) != TB_FLUID) continue;
                double dx = o->dx_c[i];
                double dV = dx * dy * dz;
                size_t id = ora_idx(o, i, j, k);
                ep += c[id] * dV * (o->yc[j] - y_ref);
                stokes += c[id] * dV;
                double v_ = v_center(o, i, j, k);
                conv += v_ * c[id] * (dx * dy * dz);
            }
        }
    }
    ep *= alpha;
    stokes *= fabs(field_vs(o)) * alpha;
    conv *= alpha;
    out[0] = ep;
    out[1] = ep_passive;
    out[2] = stokes;
    out[3] = conv;
    out[4] = ep_passive * fabs(field_vs(o));
    out[5] = conv + out[4] - stokes;
    return 0;
}

/* Conc_compute_ave_height_x (Conc.c:1524-1608) of the selected field:
 * h[i] = 1/(Ly Lz) sum over FLUID cells of column plane i of c dy dz */
int ora_ave_height_x(ora_ctx *o, double *h) {
    const double *c = o->f[TB_F_C];
    double coef = 1.0 / (o->cfg.ly * o->cfg.lz);
    for (int i = 0; i < o->NX; i++) {
        double t = 0.0;
        for (int k = 0; k < o->NZ; k++) {
            double dz = o->dz_c[k];
            for (int j = 0; j < o->NY; j++)
                if (ora_status(o, TB_Q_C, i, j, k) == TB_FLUID) t += c[ora_idx(o, i, j, k)] * o->dy_c[j] * dz;
        }
        h[i] = t * coef;
    }
    return 0;
}

/* Conc_find_front_location (Conc.c:1610-1673): from the right, the first
 * i with sum_n alpha_n h_n[i] <= limit <= sum_n alpha_n h_n[i-1]; the front
 * by linear interpolation of x_c in the height (MyMath_interpolate_quantity);
 * -1 when there is none */
double ora_front_location(ora_ctx *o, double limit) {
    int sel = o->conc_sel;
    double *h = (double *)calloc((size_t)o->NX * o->nconc, sizeof(double));
    for (int n = 0; n < o->nconc; n++) {
        ora_select_conc(o, n);
        ora_ave_height_x(o, h + (size_t)n * o->NX);
    }
    ora_select_conc(o, sel);
    double front = -1.0;
    for (int i = o->NX - 2; i > 0; i--) {
        double t0 = 0.0, t1 = 0.0;
        for (int n = 0; n < o->nconc; n++) {
            t0 += h[(size_t)n * o->NX + i - 1] * o->cs[n].p.alpha;
            t1 += h[(size_t)n * o->NX + i] * o->cs[n].p.alpha;
        }
        if (t1 <= limit && t0 >= limit) {
            front = ora_interp2(o->xc[i - 1], t0, o->xc[i], t1, limit);
            break;
        }
    }
    free(h);
    return front;
}

/* Conc_integrate_deposited_height (Conc.c:1675-1749) for the selected
 * field: at each column's interface cell, dep += dt |V_s0| w (c + c_old)/2 */
int ora_deposit_integrate(ora_ctx *o, double weight, double dt) {
    ora_conc *s = &o->cs[o->conc_sel];
    if (!s->dep) s->dep = (double *)calloc((size_t)o->NX * o->NZ, sizeof(double));
    const double *c = o->f[TB_F_C], *co = o->old[3];
    double vs = fabs(field_vs(o));
    for (int k = 0; k < o->NZ; k++)
        for (int i = 0; i < o->NX; i++) {
            int j = ora_interface_y_index(o, i, k);
            if (j < 0) continue;
            size_t id = ora_idx(o, i, j, k);
            double delta_h = dt * vs * weight * 0.5 * (c[id] + co[id]);
            s->dep[(size_t)k * o->NX + i] += delta_h;
        }
    return 0;
}

int ora_deposit_get(ora_ctx *o, double *out) {
    const ora_conc *s = &o->cs[o->conc_sel];
    size_t n = (size_t)o->NX * o->NZ;
    if (s->dep) memcpy(out, s->dep, sizeof(double) * n);
    else memset(out, 0, sizeof(double) * n);
    return 0;
}

int ora_deposit_set(ora_ctx *o, const double *in) {
    ora_conc *s = &o->cs[o->conc_sel];
    size_t n = (size_t)o->NX * o->NZ;
    if (!s->dep) s->dep = (double *)calloc(n, sizeof(double));
    memcpy(s->dep, in, sizeof(double) * n);
    return 0;
}

/* Conc_dump_particles (Conc.c:1840-1933) for the selected particle field:
 * the deposit plus the suspended column integral, dep + w sum_FLUID c dy */
int ora_deposit_dumped(ora_ctx *o, double weight, double *out) {
    const ora_conc *s = &o->cs[o->conc_sel];
    const double *c = o->f[TB_F_C];
    for (int k = 0; k < o->NZ; k++)
        for (int i = 0; i < o->NX; i++) {
            double t = 0.0;
            for (int j = 0; j < o->NY; j++)
                if (ora_status(o, TB_Q_C, i, j, k) == TB_FLUID) t += c[ora_idx(o, i, j, k)] * o->dy_c[j];
            double d = s->dep ? s->dep[(size_t)k * o->NX + i] : 0.0;
            d += t * weight;
            out[(size_t)k * o->NX + i] = d;
        }
    return 0;
}

/* Grid_get_x/y/z_index (Grid.c:1877-1967) over a quantity's coordinates,
 * with MyMath_binary_search (MyMath.c:384-398); the z search uses >= */
static int bsearch_left(const double *data, double v, int lo, int hi) {
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (v >= data[mid]) {
            if (v < data[mid + 1]) return mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return -1;
}

static int index_of(const double *xq, double x, int N, int is_z) {
    if (x < xq[0]) return ORA_NOT_FOUND;
    if (is_z ? (x >= xq[N - 1]) : (x > xq[N - 1])) return ORA_GVG_INF;
    return bsearch_left(xq, x, 0, N - 1);
}

/* MyMath_do_trilinear_inter (MyMath.c:234-360) of quantity q's field f */
static double trilinear(const ora_ctx *o, int q, const double *f, double x, double y, double z) {
    static const int dx8[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    static const int dy8[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int dz8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const double *xq, *yq, *zq;
    ora_q_coords(o, q, &xq, &yq, &zq);
    int NI = o->NX - 1, NJ = o->NY - 1, NK = o->NZ - 1;
    int i0 = index_of(xq, x, NI, 0), j0 = index_of(yq, y, NJ, 0), k0 = index_of(zq, z, NK, 1);
    if (i0 == ORA_NOT_FOUND || j0 == ORA_NOT_FOUND || k0 == ORA_NOT_FOUND) {
        i0 = ORA_MAX(i0, 0); j0 = ORA_MAX(j0, 0); k0 = ORA_MAX(k0, 0);
    }
    if (i0 == ORA_GVG_INF || j0 == ORA_GVG_INF || k0 == ORA_GVG_INF) {
        i0 = ORA_MIN(i0, NI - 2); j0 = ORA_MIN(j0, NJ - 2); k0 = ORA_MIN(k0, NK - 2);
    }
    double x0 = xq[i0], y0 = yq[j0], z0 = zq[k0];
    double x7 = xq[i0 + 1], y7 = yq[j0 + 1], z7 = zq[k0 + 1];
    double a = (x7 - x) / (x7 - x0), b = (y7 - y) / (y7 - y0), g = (z7 - z) / (z7 - z0);
    double wc[8];
    wc[0] = a * b * g;
    wc[1] = (1.0 - a) * b * g;
    wc[2] = a * (1.0 - b) * g;
    wc[3] = (1.0 - a) * (1.0 - b) * g;
    wc[4] = a * b * (1.0 - g);
    wc[5] = (1.0 - a) * b * (1.0 - g);
    wc[6] = a * (1.0 - b) * (1.0 - g);
    wc[7] = (1.0 - a) * (1.0 - b) * (1.0 - g);
    double val = 0.0;
    for (int m = 0; m < 8; m++) val += wc[m] * f[ora_idx(o, i0 + dx8[m], j0 + dy8[m], k0 + dz8[m])];
    return val;
}

static double dist3(const double *p, const double *q) {
    /* MyMath_get_point_point_distance (MyMath.c:105-111) */
    double d2 = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
    return sqrt(d2);
}

/* velocity at p and its part tangential to n; V_n is the signed normal
 * component (Extract.c:125-133) or its magnitude (Extract.c:298-303) */
static void tangential(const ora_ctx *o, const double *p, const double *n, int abs_normal, double t[3]) {
    double ui = trilinear(o, TB_Q_U, o->f[TB_F_U], p[0], p[1], p[2]);
    double vi = trilinear(o, TB_Q_V, o->f[TB_F_V], p[0], p[1], p[2]);
    double wi = trilinear(o, TB_Q_W, o->f[TB_F_W], p[0], p[1], p[2]);
    double Vn = ui * n[0] + vi * n[1] + wi * n[2];
    if (abs_normal) Vn = fabs(Vn);
    t[0] = ui - Vn * n[0];
    t[1] = vi - Vn * n[1];
    t[2] = wi - Vn * n[2];
}

/* Extract_wall_shear (Extract.c:18-196) and, with quadratic != 0,
 * Extract_wall_shear_quadratic (Extract.c:197-452) on one process (every
 * interpolation box is local): per column (i, k), x fastest, the bottom
 * shear stress tau and the tangential velocity.  At an IMMERSED interface
 * cell the velocity is interpolated at image_point1 = control + n sqrt(3)
 * dy_c[j] (Immersed.c:1135-1143) and, for the quadratic fit through the
 * control point, also at image_point2 = control + 2 n sqrt(3) dy_c[j]; at a
 * FLUID row-0 cell the cell-centre u and w over the half-cell distance. */
int ora_wall_shear(ora_ctx *o, int quadratic, double *tau, double *ut, double *vt, double *wt) {
    int *map = (int *)malloc(sizeof(int) * o->NT);
    for (size_t a = 0; a < o->NT; a++) map[a] = -1;
    for (int m = 0; m < o->nib[TB_Q_C]; m++) {
        const tb_ib_node *nd = &o->ib[TB_Q_C][m];
        map[ora_idx(o, nd->i, nd->j, nd->k)] = m;
    }
    const double a_Re = 1.0 / o->cfg.re;
    double dVt_dn = 0.0;
    for (int k = 0; k < o->NZ; k++)
        for (int i = 0; i < o->NX; i++) {
            size_t cix = (size_t)k * o->NX + i;
            int j = ora_interface_y_index(o, i, k);
            double t[3] = {0.0, 0.0, 0.0}, tv = 0.0;
            if (j != ORA_NOT_FOUND) {
                int s = ora_status(o, TB_Q_C, i, j, k);
                double d = 0.0;
                if (s == TB_IMMERSED) {
                    const tb_ib_node *nd = &o->ib[TB_Q_C][map[ora_idx(o, i, j, k)]];
                    double Dt = sqrt(3.0) * o->dy_c[nd->j];
                    double p1[3], p2[3];
                    for (int a = 0; a < 3; a++) p1[a] = nd->control[a] + nd->n[a] * Dt;
                    d = dist3(p1, nd->control);
                    tangential(o, p1, nd->n, quadratic, t);
                    if (quadratic) {
                        double Vt1 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                        for (int a = 0; a < 3; a++) p2[a] = nd->control[a] + nd->n[a] * 2.0 * Dt;
                        double d2 = dist3(p2, nd->control);
                        tangential(o, p2, nd->n, 1, t);
                        double Vt2 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                        double d0 = 0.0, Vt0 = 0.0;
                        double A[9] = {d * d, d, 1.0, d2 * d2, d2, 1.0, d0 * d0, d0, 1.0}, Ai[9];
                        ora_gauss_jordan(3, A, Ai);
                        double qa = Vt1 * Ai[0] + Vt2 * Ai[1] + Vt0 * Ai[2];
                        double qb = Vt1 * Ai[3] + Vt2 * Ai[4] + Vt0 * Ai[5];
                        dVt_dn = 2.0 * qa * d0 + qb;
                    }
                }
                if (j == 0 && s == TB_FLUID) {
                    t[0] = u_center(o, i, j, k);
                    t[1] = 0.0;
                    t[2] = w_center(o, i, j, k);
                    d = o->yc[j] - o->yv[j];
                    if (quadratic) dVt_dn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * (1.0 / d);
                }
                if (quadratic) tv = dVt_dn * a_Re;
                else tv = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) / (d * o->cfg.re);
            }
            if (tau) tau[cix] = tv;
            if (ut) ut[cix] = t[0];
            if (vt) vt[cix] = t[1];
            if (wt) wt[cix] = t[2];
        }
    free(map);
    return 0;
}
