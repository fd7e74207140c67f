/*
 * turbins_gpu.h -- C-ABI of the MI355X (gfx950) implementation of the TURBINS
 * per-timestep hot path: ENO convection, momentum/concentration Helmholtz
 * predictors with immersed-boundary rows, pressure Poisson solve and
 * projection.  Every entry point is extern "C", takes plain pointers and
 * sizes, and returns an int status:
 *      0  ok
 *     >0  finished but not converged / fallback taken (see the stats struct)
 *     <0  hard error; tb_last_error(ctx) holds the message.
 * Nothing here calls exit() (the reference aborts the process on any PETSc
 * error, definitions.h:116-123).  One host thread per context; the context owns
 * one HIP stream and all device buffers.  Host pointers are borrowed for the
 * duration of a call only.  Calls are synchronous unless named _async.
 *
 * Field layout on the host side is the reference's PETSc DA natural order:
 * (NX+1)*(NY+1)*(NZ+1) doubles, x fastest, where NX,NY,NZ are the numbers of
 * physical cells (params->NX etc.; the extra "half cell" plane is
 * Grid.c:59-62).
 *
 * Reference interfaces replaced (file:line in mmnasr/TURBINS src/):
 *   tb_eno_conv          ENO_compute_convective_terms      Include/ENO.h:6
 *                        (lifted from one grid line to whole fields, with the
 *                        wall ghosts of Velocity_*_ENO_ghost_cells
 *                        Include/Velocity.h and Conc_ENO_ghost_cells
 *                        Include/Conc.h, and the transposes
 *                        Velocity_transpose_velocities / Velocity_cell_center)
 *   tb_set_rhs           Velocity_u/v/w_set_RHS            Include/Velocity.h:15-17
 *                        Conc_set_RHS                      Include/Conc.h:30
 *   tb_helmholtz_solve   Velocity_solve                    Include/Velocity.h:36
 *                        Conc_solve                        Include/Conc.h:29
 *                        (the diagonal update of Velocity_modify_diagonal /
 *                        Conc_modify_diagonal is applied from dt)
 *   tb_poisson_rhs       Pressure_set_RHS                  Include/Pressure.h:9
 *   tb_poisson_solve     Pressure_solve                    Include/Pressure.h:11
 *   tb_project           Pressure_project_velocity         Include/Pressure.h:10
 *   tb_div_max           Pressure_compute_velocity_divergence Include/Pressure.h:12
 *   tb_setup_geometry    Grid_identify_geometry            Include/Grid.h:20
 *                        + Surface_find_exact_sdf / Surface_find_q_sdf
 *                        + Immersed_setup_q_immersed_nodes Include/Immersed.h:13
 *   tb_rk_average        Velocity_rk_average / Conc_rk_average
 *   tb_store_old         Velocity_store_old_data / Conc_store_old_data
 */
#ifndef TURBINS_GPU_H
#define TURBINS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quantities (grids of the MAC layout) */
#define TB_Q_U 0
#define TB_Q_V 1
#define TB_Q_W 2
#define TB_Q_C 3

/* node status codes: definitions.h:94-99 */
#define TB_OUTSIDE 1
#define TB_SOLID 2
#define TB_FLUID 3
#define TB_IMMERSED 5
#define TB_BOUNDARY 7
#define TB_BUFFER_SOLID 11

/* concentration field types: definitions.h:44-46 */
#define TB_SALINITY 1
#define TB_PARTICLE 2

/* field ids for upload / download */
enum {
    TB_F_U = 0, TB_F_V, TB_F_W, TB_F_C,
    TB_F_DPDX, TB_F_DPDY, TB_F_DPDZ, TB_F_PHI,
    TB_F_CONV_U, TB_F_CONV_V, TB_F_CONV_W, TB_F_CONV_C,
    TB_F_RHS_U, TB_F_RHS_V, TB_F_RHS_W, TB_F_RHS_C,
    TB_F_PRHS,              /* Poisson right-hand side  b = -div(u*)/dt      */
    TB_F_DIV,               /* |div u| per cell (G_divergence)                */
    TB_F_SDF_U, TB_F_SDF_V, TB_F_SDF_W, TB_F_SDF_C,
    TB_F_SCRATCH0, TB_F_SCRATCH1,   /* free vectors for operator tests       */
    TB_F_COUNT
};

/* Poisson modes */
#define TB_POIS_FP64_PCG 0      /* FP64 CG preconditioned by an FP64 V-cycle */
#define TB_POIS_MIXED_IR 1      /* FP32 V-cycles inside FP64 iterative refinement */
#define TB_POIS_LEARNED_IR 2    /* learned initial guess + mode 1, fallback to mode 0 */

/* Level-set reinitialisation, near-interface gradient (Levelset.c:426-456):
 * the reference drops a central difference whose stencil width |DeltaX| is
 * below an absolute 0.01 (Levelset.c:282,428-448).  That also drops interior
 * differences once 2*d_c < 0.01 (more than 200 cells over a unit length),
 * which sets D = phi/0.01 next to a flat interface and makes the pseudo-time
 * march diverge from the bottom wall.  SCALED takes min(0.01, 1.5*d_c) as the
 * guard; it equals the reference bit for bit whenever every d_c >= 0.01/1.5. */
#define TB_REINIT_SCALED 0
#define TB_REINIT_REFERENCE 1

typedef struct tb_config {
    int32_t nx, ny, nz;         /* physical cells in x, y, z (params->NX..NZ) */
    double lx, ly, lz;          /* domain lengths; origin at 0 (params->Lx..Lz) */
    double re;                  /* Reynolds number (params->Re) */
    double pe;                  /* Peclet number of the concentration field (params->Pe[0]) */
    int32_t eno_order;          /* 1..3 (params->ENO_scheme_order) */
    int32_t wall_noslip[6];     /* velocity walls x0,x1,y0,y1,z0,z1: 1 no-slip, 0 free-slip
                                   (Boundary.h:5-20 selects these at compile time) */
    double buoyancy;            /* multiplier of the interpolated c in the v right-hand
                                   side (Velocity.c:1883 ships 0.0) */
    int32_t conc_type;          /* TB_SALINITY or TB_PARTICLE (params->conc_type[0]) */
    double v_settle;            /* settling speed as used (Input.c:481 flips the sign:
                                   negative = downward) */
    int32_t sedimentation;      /* particle no-flux top wall ghost (Conc.c:1156-1181) */
    double helm_rtol;           /* Helmholtz relative residual (params->resmax) */
    int32_t helm_maxit;
    double pois_rtol;           /* Poisson relative residual target */
    int32_t pois_maxit;
    int32_t pois_mode;          /* TB_POIS_* */
    int32_t bottom_offset;      /* 1: topography is lifted by 2.001*dy_c[0] (Grid.c:913) */
    int32_t mg_pre, mg_post;    /* red-black Gauss-Seidel sweeps per V-cycle side */
    int32_t reinit_guard;       /* TB_REINIT_SCALED (0) or TB_REINIT_REFERENCE */
    int32_t reserved[7];
} tb_config;

/* One immersed-boundary node (ImmersedNode, DataTypes.h:520-546), flattened. */
typedef struct tb_ib_node {
    int32_t i, j, k;            /* im_index */
    int32_t n_fluid;            /* 8 (trilinear box) or 1 (nearest fluid node) or 0 */
    int32_t fi[8], fj[8], fk[8];/* fluid_index[] */
    double coef[8];             /* fluid_coef[] */
    double boundary_coef;
    double n[3];                /* unit normal */
    double im[3];               /* im_point */
    double control[3];          /* control_point */
    double image[3];            /* image_point */
} tb_ib_node;

typedef struct tb_solve_stats {
    int32_t iterations;         /* Krylov / refinement iterations */
    int32_t converged;          /* 1 if the target was met */
    int32_t fallback;           /* Poisson: bit 1 the mixed solve stagnated and FP64
                                   GMG-PCG finished it; bit 2 the initial guess had a
                                   larger residual than the previous phi, which FP64
                                   GMG-PCG then started from.  Helmholtz: 1 if the
                                   Gauss-Seidel sweeps stopped contracting and BiCGStab
                                   finished the solve. */
    int32_t vcycles;            /* V-cycles applied (Poisson) */
    double res0;                /* initial true relative residual (NaN: not evaluated,
                                   the Helmholtz Gauss-Seidel path) */
    double res;                 /* final true relative residual ||b-Ax||/||b|| */
    double ms;                  /* device time, milliseconds */
} tb_solve_stats;

typedef struct tb_step_stats {
    tb_solve_stats helm[2][4];  /* per RK stage, per quantity u,v,w,c */
    tb_solve_stats pois[2];     /* per RK stage */
    double div_max;             /* max |div u| after the step */
    double ms;                  /* device time of the whole step */
} tb_step_stats;

typedef struct tb_ctx tb_ctx;

/* lifecycle ------------------------------------------------------------- */
/* Grid_create (Grid.c:21) + Velocity_create (Velocity.c:19, Velocity.h:4),
 * Conc_create (Conc.c:22, Conc.h:25), Pressure_create (Pressure.c:15,
 * Pressure.h:4):
 * grid metrics, every field, the Helmholtz/Poisson operators (matrix-free)
 * and the multigrid hierarchy are set up on the device. */
tb_ctx *tb_create(const tb_config *cfg);
/* Grid_destroy (Grid.c:387), Velocity_destroy (Velocity.c:239),
 * Conc_destroy (Conc.c:221), Pressure_destroy (Pressure.c:98) */
void tb_destroy(tb_ctx *ctx);
const char *tb_last_error(tb_ctx *ctx);
int tb_device_bytes(tb_ctx *ctx, int64_t *bytes);

/* geometry: SDF reinitialisation, node classification and IB tables.
 * h_xz: bottom height at cell centres, (nz+1)*(nx+1) doubles, x fastest
 * (grid->interface_position before the offset), or NULL for no solid.
 * Replaces Grid_describe_interface + Surface_initialize +
 * Levelset_reinitialize + Surface_find_q_sdf (Grid.c:712-1121,
 * Surface.c:96-334, Levelset.c:85-709), Grid_identify_geometry
 * (Grid.c:1825-1872, Grid.h:20) and Immersed_setup_q_immersed_nodes for
 * u, v, w, c (Immersed.c:546-617, Immersed.h:13). */
int tb_setup_geometry(tb_ctx *ctx, const double *h_xz);
/* grid->G_status / G_u_status ... as uint8 (Grid.c:159-179) */
int tb_download_flags(tb_ctx *ctx, int q, uint8_t *out);
/* Immersed->N_IB (DataTypes.h:548-557) */
int tb_ib_count(tb_ctx *ctx, int q);
/* Immersed->ImmersedNodes[] in k,j,i order (Immersed.c:423-444) */
int tb_download_ib_table(tb_ctx *ctx, int q, tb_ib_node *out, int capacity);
/* pseudo-time iterations taken by Levelset_reinitialize (Levelset.c:140) */
int tb_reinit_iterations(tb_ctx *ctx);

/* grid coordinates and metrics (Grid_generate_mesh_coordinates,
 * Grid_compute_metric_coefficients, Grid.c:470-708: grid->xc ... metric_zw);
 * which: 0 xc 1 yc 2 zc 3 xu 4 yu 5 zu 6 xv 7 yv 8 zv 9 xw 10 yw 11 zw
 * 12 dx_c 13 dy_c 14 dz_c 15 metric_xc 16 metric_yc 17 metric_zc
 * 18 metric_xu 19 metric_yv 20 metric_zw.  Returns the length (nx+1, ny+1
 * or nz+1) or -1. */
int tb_grid_array(tb_ctx *ctx, int which, double *out, int capacity);

/* fields ------------------------------------------------------------------ */
/* host layout = the reference's DA natural order (padded, x fastest, FP64);
 * replaces the DAVecGetArray access to the G_data Vecs (e.g. Velocity.c:88) */
int tb_upload_field(tb_ctx *ctx, int f, const double *host);
int tb_download_field(tb_ctx *ctx, int f, double *host);

/* restart and field files ------------------------------------------------- */
/* PETSc binary vectors as the reference's PetscViewerBinary + VecView /
 * VecLoadIntoVector write and read them (Resume.c:96-139, Output.c:1610):
 * int32 class id 1211214, int32 N, N doubles, all big-endian, DA natural
 * order.  The array functions need no context (and no GPU). */
#define TB_IO_BAD_ARGUMENT (-1)
#define TB_IO_OPEN_FAILED (-2)
#define TB_IO_NOT_A_VEC (-3)   /* wrong class id, negative length or bad record */
#define TB_IO_SIZE (-4)        /* length larger than the buffer / the field */
#define TB_IO_SHORT (-5)       /* short read or write */
#define TB_IO_ZERO_STEP 1      /* warning: Resume_primary.dat's dt or dt_old reads as 0 (its
                                  six decimals lose any step below 5e-7); the values are
                                  still returned */
int tb_petsc_write_array(const char *path, const double *a, int64_t n);
/* returns N (<= capacity) or a TB_IO_ code */
int64_t tb_petsc_read_array(const char *path, double *a, int64_t capacity);
/* one field <-> one PETSc binary file (VecView / VecLoadIntoVector of a
 * G_data Vec); the file's length must equal the field's */
int tb_save_field(tb_ctx *ctx, int f, const char *path);
int tb_load_field(tb_ctx *ctx, int f, const char *path);
/* Resume_primary.dat (Resume.c:247-319): the ASCII record
 * "time=\n%lf\ndt=\n%lf\ndt_old=\n%lf\niteration=\n%d\noutput_iter=\n%d\n" */
typedef struct {
    double time, dt, dt_old;
    int32_t iter, output_iter;
} tb_resume_info;
int tb_write_resume_primary(const char *dir, const tb_resume_info *info);
/* 0, TB_IO_ZERO_STEP (record read, a step is 0) or a negative TB_IO_ code */
int tb_read_resume_primary(const char *dir, tb_resume_info *info);
/* the restart set of Resume_write_data / Resume_read_data with the files of
 * Resume_set_data (Resume.c:172-244): Resume_u, _v, _w, _dpdx, _dpdy, _dpdz,
 * _conc0 (.bin) and Resume_primary.dat in directory dir (Resume_gridc.bin,
 * written by the reference only on request, is not: the geometry is set up
 * again from the topography) */
int tb_write_resume(tb_ctx *ctx, const char *dir, const tb_resume_info *info);
/* 0, TB_IO_ZERO_STEP (fields read; dt or dt_old is 0) or -1 */
int tb_read_resume(tb_ctx *ctx, const char *dir, tb_resume_info *info);

/* per-operation mirrors of the reference entry points ---------------------- */
/* advective term q -> TB_F_CONV_q for the whole field: replaces the
 * per-line ENO_compute_convective_terms (ENO.c:103-245, ENO.h:6) together
 * with its gathers, the wall ghosts (Velocity.c:2124-2801,
 * Conc.c:1021-1220), Velocity_transpose_velocities (Velocity.c:3341-3663)
 * and Velocity_cell_center (Velocity.c:1476-1602) */
int tb_eno_conv(tb_ctx *ctx, int q);
/* Velocity_u/v/w_set_RHS (Velocity.c:1607-2043, Velocity.h:15-17),
 * Conc_set_RHS (Conc.c:864-992, Conc.h:30) */
int tb_set_rhs(tb_ctx *ctx, int q, double dt);
/* tb_eno_conv(ctx, -1) followed by tb_set_rhs(ctx, q, dt) for q = 0..3, in
 * one pass of the ENO kernels (same values, bit for bit; the convective terms
 * are also stored): ENO_compute_convective_terms (ENO.c:103-245) and
 * Velocity_u/v/w_set_RHS (Velocity.c:1607-2043) / Conc_set_RHS
 * (Conc.c:864-992) of one RK stage; tb_step uses it */
int tb_eno_conv_rhs(tb_ctx *ctx, double dt);
/* y = A_q x with the rows of Velocity_*_form_LHS_matrix
 * (Velocity.c:449-1471) / Conc_form_LHS_matrix (Conc.c:390-730) after
 * Velocity_modify_diagonal / Conc_modify_diagonal for dt
 * (Velocity.c:374-443, Conc.c:293-354) */
int tb_helmholtz_apply(tb_ctx *ctx, int q, double dt, int f_in, int f_out);
/* Velocity_solve (Velocity.c:306-324, Velocity.h:36) / Conc_solve
 * (Conc.c:840-859, Conc.h:29): previous field as the initial guess, stop at
 * ||b - A x|| <= helm_rtol ||b|| on the true residual, as the reference's KSP
 * (Solver.c:39-86, 269-322).  Method per tb_set_helmholtz_method: red-black
 * Gauss-Seidel sweep pairs with BiCGStab taking over if they stall
 * (TB_HELM_GS, default), or the reference's unpreconditioned BiCGStab
 * (TB_HELM_BICGSTAB). */
int tb_helmholtz_solve(tb_ctx *ctx, int q, double dt, tb_solve_stats *st);
/* The smoother of TB_HELM_GS by itself (no reference counterpart; exposed
 * for parity tests): f_x <- `pairs` red-black sweep pairs of
 * A_q(dt) x = f_b (red = (i+j+k) even first; IMMERSED rows skipped, then
 * one Jacobi update of the IMMERSED rows per pair), in place. */
int tb_helmholtz_gs_sweeps(tb_ctx *ctx, int q, double dt, int f_b, int f_x, int pairs);
/* Pressure_set_RHS (Pressure.c:171-276, Pressure.h:9) */
int tb_poisson_rhs(tb_ctx *ctx, double dt);
/* y = A_p x, rows of Pressure_form_LHS_matrix (Pressure.c:691-936) */
int tb_poisson_apply(tb_ctx *ctx, int f_in, int f_out);
/* Pressure_solve (Pressure.c:443-458, Pressure.h:11), which returns the
 * iteration count; here st->iterations.  Krylov/multigrid method chosen by
 * tb_set_poisson_mode (the reference uses GMRES + BoomerAMG, Solver.c:90-265) */
int tb_poisson_solve(tb_ctx *ctx, tb_solve_stats *st);
/* Pressure_solve from a caller-supplied initial guess in field f_guess (no
 * reference counterpart: the reference starts from the previous phi,
 * Solver.c:256; this is the entry point for a guess predicted outside the
 * library).  The guess is pinned and kept when its residual is not larger
 * than that of the previous phi; otherwise the previous phi is restored and
 * FP64 GMG-PCG solves from it (st->fallback bit 2).  Mode as set by
 * tb_set_poisson_mode; f_guess is neither TB_F_PHI nor TB_F_PRHS. */
int tb_poisson_solve_guess(tb_ctx *ctx, int f_guess, tb_solve_stats *st);
/* Pressure_compute_real_pressure (Pressure.c:590-685, called before pressure
 * output, Output.c:2198-2204): TB_F_PRHS = -div(grad p) from TB_F_DPDX/Y/Z
 * (0 on BOUNDARY cells), then the Poisson solve into TB_F_PHI, which becomes
 * the next solve's plain initial guess as in the reference.  This output
 * solve leaves the learned mode's solution history, its Galerkin matrix, the
 * snapshot store and the network's judgement as they were. */
int tb_real_pressure_rhs(tb_ctx *ctx);
/* diagnostics: Extract_total_kinetic_energy (Extract.c:803-884) and
 * Conc_compute_total_suspended_mass (Conc.c:1471-1512), both over FLUID
 * cells of the concentration grid, in one pass; either pointer may be NULL.
 * Summation order differs from the reference's serial loop (FP64, relative
 * differences of order 1e-15 per 1e6 cells). */
int tb_diagnostics(tb_ctx *ctx, double *kinetic_energy, double *suspended_mass);
/* Extract_Stokes_dissipation_rate (Extract.c:886-950): sum of c dV over FLUID
 * cells times |settling speed| times conc_alpha (params->Conc_alpha[i]) for
 * a particle concentration, 0 for salinity */
int tb_stokes_dissipation(tb_ctx *ctx, double conc_alpha, double *rate);
/* Conc_compute_sedimentation_rate (Conc.c:2098-2157) and Conc_bottom_sed_rate
 * (Conc.c:2158-2229): per (i, k) column |v_s| c dx dz at the column's first
 * IMMERSED cell, or at row 0 when it is FLUID (Grid_set_interface_y_index,
 * Grid.c:1972-2009), 0 elsewhere; bottom receives the (NZ+1) x (NX+1) column
 * values (x fastest, may be NULL), total their sum in the reference's order;
 * 0 for salinity */
int tb_sedimentation_rate(tb_ctx *ctx, double *total, double *bottom);
int tb_real_pressure(tb_ctx *ctx, tb_solve_stats *st);
/* Pressure_project_velocity (Pressure.c:280-439, Pressure.h:10) */
int tb_project(tb_ctx *ctx, double dt);
/* Pressure_compute_velocity_divergence (Pressure.c:462-586, Pressure.h:12);
 * the same pass scans u, v, w and c for NaN / Inf on the non-BOUNDARY cells
 * (Debugger_check_nan, Debugger.c:286-324) and returns -1 with the cell in
 * ijk and in tb_last_error when it finds one (tb_step then fails too) */
int tb_div_max(tb_ctx *ctx, double *div_max, int32_t *ijk);
/* Allocation-free divergence check.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * kernels as tb_div_max on caller-owned device buffers, no context: |div u|
 * per cell into div (device, may be NULL), the maximum into *out_max and its
 * cell index (k*(ny+1) + j)*(nx+1) + i into *out_index (both device
 * pointers), enqueued on `stream` (a hipStream_t; NULL: the default stream)
 * and not waited for.  A NaN / Inf in u, v, w or c on a non-BOUNDARY cell
 * makes *out_max +infinity with that cell's index.  Fields are FP64 in the
 * layout of the library's fields with row pitch box->pitch >= nx+1 elements
 * (x fastest, then y, z); flags_c are the concentration grid's status bytes
 * in the same layout; mxc, myc, mzc are the metric arrays (Grid.c:595-708)
 * of nx+1, ny+1, nz+1 doubles on the device.  Workspace: a 32 KiB device
 * array inside the library (one per GPU, held until the process ends); calls
 * that overlap on one GPU race for it.  Boxes of more than 2^31 points are
 * rejected.  Calls no hipMalloc / hipFree.  Returns 0 or a TB_NOALLOC_
 * code. */
typedef struct tb_box {
    int32_t nx, ny, nz;         /* physical cells */
    int32_t pitch;              /* row pitch of every field, in elements */
    double deta, dxi, dphi;     /* computational spacings 1/nx, 1/ny, 1/nz (grid->dEta ...) */
} tb_box;
#define TB_NOALLOC_BAD_ARGUMENT (-1)
#define TB_NOALLOC_TOO_LARGE (-2)     /* more than 2^31 points */
#define TB_NOALLOC_NO_SCRATCH (-3)    /* no current device, or its workspace address unavailable */
#define TB_NOALLOC_LAUNCH (-4)        /* kernel launch failed */
int tb_div_max_noalloc(const tb_box *box, const double *mxc, const double *myc, const double *mzc,
                       const double *u, const double *v, const double *w, const double *c,
                       const uint8_t *flags_c, double *div, double *out_max, int64_t *out_index,
                       void *stream);
/* Velocity_store_old_data + Conc_store_old_data (Velocity.c:297-302,
 * Velocity.h:6; Conc.c:997-1003, Conc.h:31) for u, v, w, c */
int tb_store_old(tb_ctx *ctx);
/* Velocity_rk_average + Conc_rk_average (Velocity.c:2047-2055,
 * Velocity.h:19; Conc.c:1007-1017, Conc.h:19): q = old_frac*q_old + new_frac*q */
int tb_rk_average(tb_ctx *ctx, double old_frac, double new_frac);

/* CFL time step of the host driver (the reference's parameters cfl_method,
 * alpha = 0.5 and default_dt, Input.c:421-423; its rule lived in the
 * unpublished driver, README.md:26): over the FLUID cells of the
 * concentration grid, with the cell-centre velocities of Velocity_cell_center
 * (Velocity.c:1476-1602), rate = |u_c|/dx + |v_c|/dy + |w_c|/dz and
 * dt = min(dt_max, alpha / max rate) (dt_max when the velocity is 0).
 * rate (may be NULL) receives the maximum; NaN / Inf velocities fail.  The
 * Helmholtz rows take dt per call, which is the diagonal the reference's
 * Velocity_modify_diagonal(dt, dt_old) accumulates (+1/dt - 1/dt_old). */
int tb_cfl_dt(tb_ctx *ctx, double alpha, double dt_max, double *dt, double *rate);

/* Concentration fields (params->NConc and the per-field arrays Pe, V_s0,
 * Conc_alpha, Phi, conc_type read in Input.c; one Concentration object per
 * field, Conc_create, Conc.c:22-217).  A context starts with one field described by
 * tb_config (pe, conc_type, v_settle). */
#define TB_NCONC_MAX 4
typedef struct tb_conc_field {
    int32_t type;               /* TB_SALINITY or TB_PARTICLE (params->conc_type[i]) */
    int32_t reserved;
    double pe;                  /* Peclet number (params->Pe[i]) */
    double v_settle;            /* settling speed as used, negative = downward
                                   (params->V_s0[i] after Input.c:480 flips it) */
    double alpha;               /* weight in the total concentration (params->Conc_alpha[i]) */
    double phi_max;             /* particle volume fraction at c = 1 (params->Phi[i]) */
} tb_conc_field;
/* Sets nconc (1..TB_NCONC_MAX) fields; fields[0] replaces tb_config's pe,
 * conc_type and v_settle.  Fields 1.. are allocated (zero) or freed as
 * nconc grows or shrinks; field 0 keeps its data.  hindered != 0 is
 * params->hindered_settling (Input.c:369): a particle field is then advected
 * with v + V_s0 (1 - c phi_max)^4 on FLUID cells and v elsewhere
 * (Conc_compute_particle_settling_speed, Conc_settling_speed_function and
 * Conc_set_particle_v_velocity, Conc.c:734-836; the power is taken as two
 * squarings, within 2 ulp of pow()), instead of v + V_s0 everywhere.  The
 * no-flux top wall of a sedimenting field keeps |V_s0| (Conc.c:1156-1181).
 * Selects field 0. */
int tb_set_conc_fields(tb_ctx *ctx, int nconc, const tb_conc_field *fields, int hindered);
/* Makes field n the one that the TB_Q_C operators (tb_eno_conv, tb_set_rhs,
 * tb_helmholtz_*, tb_diagnostics' mass, tb_stokes_dissipation,
 * tb_sedimentation_rate) and the transfers of TB_F_C, TB_F_CONV_C and
 * TB_F_RHS_C act on: the reference passes c[iconc] to each Conc_* call. */
int tb_select_conc(tb_ctx *ctx, int n);
/* Conc_compute_total_concentration (Conc.c:1291-1314): with more than one
 * field, c_total = alpha_0 c_0, then c_total += alpha_i c_i for i = 1..; the
 * v right-hand side then takes its buoyancy term from c_total instead of c_0
 * (Velocity.c:1798-1815).  A no-op for one field.  tb_step calls it at the
 * start of each stage. */
int tb_total_concentration(tb_ctx *ctx);

/* Output-stage diagnostics (Output_flow_properties, Output.c:2062-2216).
 * Column arrays are (nz+1) x (nx+1) doubles, x fastest; an output pointer
 * may be NULL.  Sums over cells run in a fixed order of their own (FP64,
 * relative differences from the reference's serial loop of order 1e-15 per
 * 1e6 cells); column values and sums over columns follow the reference's
 * operations and order. */
/* Extract_viscous_dissipation_rate (Extract.c:453-725), without inflow or
 * outflow: (1/Re) sum over FLUID cells of u lap(u) dV_u + v lap(v) dV_v +
 * w lap(w) dV_w, with the wall ghosts the reference builds from Boundary.h
 * (no-slip -u, free-slip +u; the wall-normal component 0) */
int tb_viscous_dissipation(tb_ctx *ctx, double *rate);
/* Extract_kinetic_potential_conversion_rate (Extract.c:727-802): sum over
 * FLUID cells of -v_c c_total dV (the total of tb_total_concentration,
 * recomputed here, with several fields; field 0 with one) */
int tb_kinetic_potential_conversion(tb_ctx *ctx, double *rate);
/* Conc_compute_potential_energies (Conc.c:1934-2058) of the selected field */
typedef struct tb_energies {
    double ep;                  /* alpha sum c dV (y_c - y_v[0]) over FLUID cells */
    double ep_passive;          /* particles: sum deposit dx dz (y_c[j_s] - y_v[0]) */
    double stokes;              /* |V_s0| alpha sum c dV (particles) */
    double conversion;          /* alpha sum v_c c dV */
    double ep_passive_rate;     /* ep_passive |V_s0| */
    double ep_rate;             /* conversion + ep_passive_rate - stokes */
} tb_energies;
int tb_potential_energies(tb_ctx *ctx, tb_energies *out);
/* Extract_wall_shear (Extract.c:18-196), or with quadratic != 0
 * Extract_wall_shear_quadratic (Extract.c:197-452): per (i, k) column of
 * the bottom (the interface cell of Grid_set_interface_y_index,
 * Grid.c:1972-2009) the shear stress and the tangential velocity.  At an
 * IMMERSED cell the velocity is interpolated trilinearly
 * (MyMath_do_trilinear_inter, MyMath.c:234-360) at image_point1 = control
 * point + n sqrt(3) dy_c (Immersed.c:1135-1143), and for the quadratic fit
 * through the control point also at image_point2 (twice as far); at a FLUID
 * cell of row 0 the cell-centre u and w over the half-cell height.  Columns
 * without an interface cell give 0. */
int tb_wall_shear(tb_ctx *ctx, int quadratic, double *tau, double *u_tan, double *v_tan, double *w_tan);
/* Conc_compute_ave_height_x (Conc.c:1524-1608) of the selected field: nx+1
 * values h[i] = 1/(Ly Lz) sum over the FLUID cells of plane i of c dy dz */
int tb_ave_height_x(tb_ctx *ctx, double *h);
/* Conc_find_front_location (Conc.c:1610-1673, front_limit 0.01 at
 * Output.c:2139): the front from the span-averaged heights of every field
 * weighted by alpha; -1 when no crossing is found */
int tb_front_location(tb_ctx *ctx, double front_limit, double *x_front);
/* Conc_integrate_deposited_height (Conc.c:1675-1749) of the selected field,
 * after a step: at each column's interface cell the deposit grows by
 * dt |V_s0| weight (c + c_old)/2, c_old the field at the start of the step
 * (0 growth for salinity) */
int tb_deposit_integrate(tb_ctx *ctx, double weight, double dt);
/* the selected field's deposit (W_deposit_height, Conc.c:1751-1793), and
 * setting it (a restart, or 0 to reset) */
int tb_deposit_height(tb_ctx *ctx, double *h);
int tb_deposit_set(tb_ctx *ctx, const double *h);
/* Conc_dump_particles (Conc.c:1840-1933) of the selected field: the deposit
 * plus weight times the column integral of the suspended c dy */
int tb_deposit_dumped(tb_ctx *ctx, double weight, double *h);

/* ---- open x boundaries (Inflow.c, Outflow.c) ----
 * A closed box by default.  With inflow the x = 0 face is open: u there
 * takes the inflow plane (u->G_inflow) and every concentration field its own
 * (c->G_inflow).  With outflow the x = Lx face is open: u, v, w and every
 * concentration field take there the planes of the convective boundary
 * condition (G_outflow).  The planes enter the x-line ENO ghosts
 * (Velocity.c:2193-2232, 2442-2444, 2669-2671; Conc.c:1099-1122), the
 * transposes (Velocity.c:3491-3505, 3589-3603), the cell-centre velocities
 * (Velocity.c:1544-1554), the right-hand sides (Velocity.c:1691-1721,
 * 1885-1897, 2011-2023; Conc.c:938-963), the divergence (Pressure.c:242-249,
 * 541-548) and the viscous dissipation (Extract.c:577-645); the open faces
 * drop their wall rule from the Helmholtz rows (Velocity.c:989, 1336;
 * Conc.c:582-607).  A plane holds NY*NZ doubles indexed [k][j], j fastest,
 * over the padded point counts NY = ny+1, NZ = nz+1. */
typedef struct tb_open_bc {
    int32_t inflow;             /* params->inflow (Input.c:371) */
    int32_t outflow;            /* params->outflow (Input.c:372) */
    int32_t transient_inflow;   /* params->transient_inflow (Input.c:446) */
    int32_t reserved;
    double u_conv;              /* params->U, the convective outflow speed (Input.c:424) */
    double time_c;              /* params->time_c (Input.c:447) */
} tb_open_bc;
enum { TB_PLANE_INFLOW = 0, TB_PLANE_OUTFLOW = 1 };
/* Opens or closes the x faces; the planes start at zero (calloc'd as
 * Memory_allocate_2D_double does), the influx at zero. */
int tb_set_open_boundaries(tb_ctx *ctx, const tb_open_bc *bc);
/* Inflow_u_velocity_profile (Inflow.c:13-158) and Inflow_conc_profile
 * (Inflow.c:161-260) at time t: 1 times the transient factor
 * 1 - exp(-(t + 0.01)/time_c) on the FLUID u nodes of i = 1 and on the
 * non-SOLID cells of i = 0 of every concentration field (other entries keep
 * their values); *influx = sum of u_in dy dz over those u nodes
 * (u->W_influx), or null */
int tb_inflow_profile(tb_ctx *ctx, double time, double *influx);
/* A plane of quantity q (TB_Q_U .. TB_Q_C, the selected concentration
 * field for TB_Q_C; inflow planes exist for u and c only).  Setting the u
 * inflow plane recomputes the influx as tb_inflow_profile does. */
int tb_set_boundary_plane(tb_ctx *ctx, int which, int q, const double *plane);
int tb_get_boundary_plane(tb_ctx *ctx, int which, int q, double *plane);
int tb_get_influx(tb_ctx *ctx, double *influx);
/* Outflow_impose_convective_boundary (Outflow.c:12-101, 104-159): every
 * outflow plane of u, v, w and each concentration field advances as
 * out = (1 - c) out + c q[NI-1] with c = dt U / dx at the rows whose node
 * at i = NI-1 is FLUID or IMMERSED; then the u plane takes the correction
 * (influx - outflux) / area over those rows so that its outflux equals the
 * influx (Outflow_update_u_velocity_to_conserve_mass, Outflow.c:163-304).
 * tb_step calls it at the start of each stage; tb_store_old and
 * tb_rk_average carry the outflow planes with the fields
 * (Velocity_store_old_outflow / Velocity_rk_average_outflow,
 * Velocity.c:2058-2120; Conc.c:1223-1289). */
int tb_outflow_convective(tb_ctx *ctx, double dt);

/* composite SSP-RK2 step: the per-step sequence of the unpublished driver
 * (README.md:26), reconstructed from the TVD_RK2 scheme (definitions.h:148-149);
 * every concentration field is stepped */
int tb_step(tb_ctx *ctx, double dt, tb_step_stats *st);

/* solver controls (no reference counterpart: the reference fixes these in
 * Solver.c / Input.c at build time) */
int tb_set_poisson_mode(tb_ctx *ctx, int mode, double rtol);
int tb_set_helmholtz_rtol(tb_ctx *ctx, double rtol);
/* Helmholtz solver: TB_HELM_GS (default) red-black Gauss-Seidel sweeps with
 * BiCGStab taking over if they stop contracting; TB_HELM_BICGSTAB the
 * reference's method alone (KSPIBCGS, Solver.c:39-86, 269-322).  Both stop
 * at ||b - A x|| <= rtol ||b||. */
#define TB_HELM_GS 0
#define TB_HELM_BICGSTAB 1
int tb_set_helmholtz_method(tb_ctx *ctx, int method);
/* mixed Poisson mode: smallest residual reduction one FP32 inner solve is
 * asked for (default 1e-5) and its iteration cap (default 40); a cap of 0
 * leaves the refinement passes without FP32 work, so every mixed solve
 * hands over to FP64 GMG-PCG at once (st->fallback bit 1) */
int tb_set_mixed_inner(tb_ctx *ctx, double reduction_floor, int maxit);
/* learned mode: number of previous solutions (1..10, default 8) whose
 * A-norm-optimal combination, with the network's direction when one is
 * attached and its gate open, is the initial guess (the reference starts
 * from the previous pressure alone, Solver.c:256) */
int tb_set_guess_history(tb_ctx *ctx, int h);
/* sweeps of the finest multigrid level per V-cycle side, FP32 and FP64
 * cycles apart (0: mg_pre / mg_post of the config, which the coarser levels
 * keep; 1..8 otherwise).  No reference counterpart: hypre's BoomerAMG
 * settings (Solver.c:90-265) are fixed there. */
int tb_set_mg_finest_sweeps(tb_ctx *ctx, int pre32, int post32, int pre64, int post64);
/* multigrid launch structure: 1 (default) the levels of at most 2^18 cells
 * between the finest and the LDS-resident coarse cycle run each V-cycle
 * side as one launch (all sweep pairs and, after the coarse levels, the
 * prolongation, per box in LDS), and the FP32 fused-pair levels coarsened
 * in y and z form the prolongation inside their first post-smoothing pair;
 * 0 one launch per half sweep and a separate prolongation.  Both give the
 * same bits.  No reference counterpart (hypre's BoomerAMG,
 * Solver.c:90-265). */
int tb_set_mg_smoother(tb_ctx *ctx, int blocked);
/* multigrid V-cycles replayed from HIP graphs (default 1) or launched
 * kernel by kernel (0); results are identical */
int tb_set_graphs(tb_ctx *ctx, int on);

/* learned initial guess of the Poisson solve (no reference counterpart: the
 * reference starts GMRES from the previous phi, Solver.c:256).  A 3-level
 * 3-D conv encoder-decoder (16/32/64 channels, 3x3x3, BF16 weights, FP32
 * accumulation) maps (b_n, phi_{n-1}, phi_{n-2}) to a correction of
 * phi_{n-1}.  The network lives on the device of its creator; a context
 * uses it in TB_POIS_LEARNED_IR mode once attached.  Cell counts of every
 * grid it runs on must be multiples of 4.                                  */
typedef struct tb_net tb_net;
tb_net *tb_net_create(uint64_t seed);            /* He-uniform initial weights */
void tb_net_destroy(tb_net *net);
const char *tb_net_last_error(tb_net *net);
int64_t tb_net_param_count(void);
int tb_net_get_weights(tb_net *net, float *host, int64_t n);
int tb_net_set_weights(tb_net *net, const float *host, int64_t n);
int tb_net_attach(tb_ctx *ctx, tb_net *net);     /* NULL detaches */
/* The learned mode starts each solve from the Galerkin (A-norm optimal)
 * combination of phi_{n-1}, phi_{n-2} and the network's correction.  The
 * first 4 solves with the network estimate the iterations its direction
 * saves (energy of the guesses with and without it); if those iterations
 * cost less time than the inference, the context stops calling the network
 * (tb_net_attach re-opens).  This reports that judgement.                  */
int tb_net_gate(tb_ctx *ctx, int *open, int *trials, double *saved_iters_per_solve, double *net_ms_per_call);
/* guess for the current PRHS and solution history into field f_out */
int tb_net_predict_field(tb_ctx *ctx, int f_out);
/* training data: the solver's own (b_n, phi_{n-1}, phi_{n-2}) -> phi_n
 * pairs, taken after each solve of a context on an nx*ny*nz cell grid */
int tb_net_reserve(tb_net *net, int capacity, int nx, int ny, int nz);
int tb_net_add_snapshot(tb_net *net, tb_ctx *ctx);
/* from now on every Poisson solve of ctx adds its snapshot (until the store
 * is full); NULL stops */
int tb_net_record(tb_ctx *ctx, tb_net *net);
int tb_net_snapshot_count(tb_net *net);
/* Adam on random (cx,cy,cz) crops, 'batch' crops per step; mean-square
 * loss of the first and the last step */
int tb_net_train(tb_net *net, int steps, int batch, int cx, int cy, int cz, double lr, uint64_t seed,
                 double *loss_first, double *loss_last);
int tb_net_eval_snapshots(tb_net *net, int first, int count, double *mean_loss);
/* host-buffer forward / loss gradient on an nx*ny*nz grid (in3: x fastest,
 * 3 channels per voxel), for checks against a CPU restatement */
int tb_net_forward(tb_net *net, const float *in3, int nx, int ny, int nz, float *out);
int tb_net_gradient(tb_net *net, const float *in3, const float *target, int nx, int ny, int nz,
                    double *loss, float *grad);

/* live per-kernel timing with HIP events on the context's stream (no
 * reference counterpart; replaces PETSc's -log_summary for this path) */
enum {
    TB_K_ENO_MOM = 0, TB_K_ENO_CONC, TB_K_HELM_OP, TB_K_BICG_UPD, TB_K_POIS_OP,
    TB_K_RBGS0, TB_K_RBGS0_F32, TB_K_PCG_UPD, TB_K_PROJECT, TB_K_RHS, TB_K_NET,
    TB_K_CG32_OP, TB_K_CG32_UPD, TB_K_HELM_GS,
    TB_K_MG_LEVELS,      /* sweeps of the multigrid levels below the finest */
    TB_K_MG_TRANSFER,    /* residual + restriction and prolongation, every level */
    TB_K_MG_SMALL,       /* the one-workgroup cycle of the small levels, or the coarse sweeps */
    TB_K_GUESS,          /* the Galerkin guess's dot products and combination */
    TB_K_IB,             /* the immersed rows' Jacobi updates and scaled applications */
    TB_K_COUNT
};
int tb_profile_enable(tb_ctx *ctx, int on);
int tb_profile_read(tb_ctx *ctx, int kernel, double *total_ms, int64_t *calls);

#ifdef __cplusplus
}
#endif
#endif
