/*
 * tb_oracle.h -- CPU oracle for the TURBINS per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithms (mmnasr/TURBINS src/ C files, cited per function in the
 * sources).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline -- never as part of the product path.
 *
 * Layout: every 3-D field is (NX)*(NY)*(NZ) doubles with NX = nx+1 etc.,
 * x fastest (PETSc DA natural order, Grid.c:59-62).
 */
#ifndef TB_ORACLE_H
#define TB_ORACLE_H

#include <stdint.h>
#include "../include/turbins_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

ora_ctx *ora_create(const tb_config *cfg);
void ora_destroy(ora_ctx *o);

/* grid coordinates / metrics (Grid.c:470-708).  which: 0 xc 1 yc 2 zc
 * 3 xu 4 yu 5 zu 6 xv 7 yv 8 zv 9 xw 10 yw 11 zw 12 dx_c 13 dy_c 14 dz_c
 * 15 mxc 16 myc 17 mzc 18 mxu 19 myv 20 mzw */
int ora_get_grid_array(ora_ctx *o, int which, double *out);

int ora_setup_geometry(ora_ctx *o, const double *h_xz);
int ora_reinit_iterations(ora_ctx *o);
int ora_set_geometry(ora_ctx *o, int q, const uint8_t *flags, const tb_ib_node *ib, int n);
int ora_get_flags(ora_ctx *o, int q, uint8_t *out);
int ora_ib_count(ora_ctx *o, int q);
int ora_get_ib(ora_ctx *o, int q, tb_ib_node *out, int cap);

int ora_set_field(ora_ctx *o, int f, const double *in);
int ora_get_field(ora_ctx *o, int f, double *out);

/* ENO_compute_convective_terms on one line (ENO.c:103-245) */
void ora_eno_line(int order, int n, const double *D0, const double *pos,
                  const double *vel, double *vdqdx);
/* 8x8 Gauss-Jordan inverse (MyMath.c:117-209), row-major in/out */
void ora_gauss_jordan(int n, const double *a, double *a_inv);

int ora_eno_conv(ora_ctx *o, int q);
int ora_set_rhs(ora_ctx *o, int q, double dt);
int ora_helmholtz_apply(ora_ctx *o, int q, double dt, const double *x, double *y);
int ora_helmholtz_gs_sweeps(ora_ctx *o, int q, double dt, const double *b, double *x, int pairs);
int ora_helmholtz_matrix(ora_ctx *o, int q, double dt, int32_t *cnt, int64_t *col, double *val);
int ora_poisson_matrix(ora_ctx *o, int32_t *cnt, int64_t *col, double *val);
int ora_helmholtz_solve(ora_ctx *o, int q, double dt, double rtol, int maxit,
                        tb_solve_stats *st);
int ora_poisson_rhs(ora_ctx *o, double dt);
int ora_real_pressure_rhs(ora_ctx *o);
int ora_diagnostics(ora_ctx *o, double *ke, double *mass);
double ora_stokes_dissipation(ora_ctx *o, int particle, double v_settl, double conc_alpha);
double ora_sedimentation_rate(ora_ctx *o, int particle, double v_settl, double *bottom);
int ora_poisson_apply(ora_ctx *o, const double *x, double *y);
int ora_poisson_solve(ora_ctx *o, double rtol, int maxit, tb_solve_stats *st);
int ora_project(ora_ctx *o, double dt);
double ora_div_max(ora_ctx *o, int32_t *ijk);
int ora_store_old(ora_ctx *o);
int ora_rk_average(ora_ctx *o, double old_frac, double new_frac);
int ora_step(ora_ctx *o, double dt, tb_step_stats *st);
/* concentration fields: tb_set_conc_fields / tb_select_conc /
 * tb_total_concentration (Conc.c:22-217, 734-836, 1291-1314) */
int ora_set_conc_fields(ora_ctx *o, int nconc, const tb_conc_field *fields, int hindered);
int ora_select_conc(ora_ctx *o, int n);
int ora_total_concentration(ora_ctx *o);
/* Conc_settling_speed_function (Conc.c:815-836) as restated: V_s0 (1 - c phi)^4 */
double ora_hindered_speed(double conc, double phi_max, double v_s0);
double ora_cfl_dt(ora_ctx *o, double alpha, double dt_max, double *rate);
/* open x faces: tb_set_open_boundaries, tb_inflow_profile,
 * tb_set/get_boundary_plane, tb_get_influx, tb_outflow_convective
 * (Inflow.c:13-260, Outflow.c:12-304) */
int ora_set_open_boundaries(ora_ctx *o, const tb_open_bc *bc);
int ora_inflow_profile(ora_ctx *o, double time, double *influx);
int ora_set_boundary_plane(ora_ctx *o, int which, int q, const double *plane);
int ora_get_boundary_plane(ora_ctx *o, int which, int q, double *plane);
double ora_get_influx(ora_ctx *o);
int ora_outflow_convective(ora_ctx *o, double dt);

/* output-stage diagnostics (ora_diag.c); per-column arrays are
 * (NZ+1) x (NX+1), x fastest */
int ora_interface_y_index(const ora_ctx *o, int i, int k);
double ora_viscous_dissipation(ora_ctx *o);
double ora_kinetic_potential_conversion(ora_ctx *o);
int ora_potential_energies(ora_ctx *o, double *out);
int ora_ave_height_x(ora_ctx *o, double *h);
double ora_front_location(ora_ctx *o, double limit);
int ora_deposit_integrate(ora_ctx *o, double weight, double dt);
int ora_deposit_get(ora_ctx *o, double *out);
int ora_deposit_set(ora_ctx *o, const double *in);
int ora_deposit_dumped(ora_ctx *o, double weight, double *out);
int ora_wall_shear(ora_ctx *o, int quadratic, double *tau, double *ut, double *vt, double *wt);

/* GMG hierarchy description (for tests): returns number of levels and fills
 * dims[3*l..3*l+2] */
int ora_gmg_levels(ora_ctx *o, int32_t *dims, int cap);
int ora_gmg_vcycle(ora_ctx *o, const double *r_compact, double *z_compact);

/* learned-guess network (ora_net.c): parameters w (ora_net_param_count
 * doubles, the device layout), in3 [nz][ny][nx][3], out/target [nz][ny][nx];
 * round = 1 reproduces the device's BF16/FP32 rounding points */
int64_t ora_net_param_count(void);
int ora_net_forward(const double *w, const double *in3, int nx, int ny, int nz, int round, double *out);
int ora_net_gradient(const double *w, const double *in3, const double *target, int nx, int ny, int nz, int round,
                     double *loss, double *grad);

#ifdef __cplusplus
}
#endif
#endif
