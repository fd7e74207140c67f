/* Internal state of the CPU oracle (test infrastructure only). */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include "tb_oracle.h"

#define ORA_NOT_FOUND (-1)
#define ORA_GVG_INF 10000000          /* definitions.h:105 */
#define ORA_IBM_MAX 8                 /* definitions.h:102 */
#define ORA_DROP_TOL 1e-8             /* LSYS_DROP_TOLERANCE, definitions.h:152 */
#define ORA_DIRICHLET 0
#define ORA_NEUMANN 1

#define ORA_MAX(a, b) ((a) < (b) ? (b) : (a))
#define ORA_MIN(a, b) ((a) > (b) ? (b) : (a))

/* sparse matrix with at most 9 entries per row, columns sorted ascending
 * (PETSc AIJ keeps each row sorted, so MatMult sums in column order). */
typedef struct {
    size_t n;
    int32_t *cnt;
    int64_t *col;     /* n*9 */
    double *val;      /* n*9 */
} ora_mat;

typedef struct {
    int nx, ny, nz;           /* cells */
    double hx, hy, hz;
    int cx, cy, cz;           /* coarsened towards the next level */
    double *x, *b, *r;        /* compact cell arrays */
} ora_level;

/* one concentration field (Conc_create, Conc.c:22-217) */
typedef struct {
    double *c, *conv, *rhs, *old;
    double *dep;              /* deposited height per (i, k) column, x fastest */
    tb_conc_field p;
    double *cin, *cout, *cout_old;   /* open-boundary planes (G_inflow, G_outflow, G_outflow_old) */
} ora_conc;

struct ora_ctx {
    tb_config cfg;
    int ncx, ncy, ncz;
    int NX, NY, NZ;
    size_t NT;
    double dEta, dXi, dPhi, dx, dy, dz;
    double *xc, *yc, *zc, *xu, *yu, *zu, *xv, *yv, *zv, *xw, *yw, *zw;
    double *dx_c, *dy_c, *dz_c;
    double *mxc, *myc, *mzc, *mxu, *myv, *mzw;

    int has_geometry;
    uint8_t *st[4];
    int nib[4];
    tb_ib_node *ib[4];
    int reinit_its;

    double *f[TB_F_COUNT];
    double *old[4];
    /* concentration fields; the selected one is viewed through f[TB_F_C],
     * f[TB_F_CONV_C], f[TB_F_RHS_C], old[3] and cfg.pe / conc_type / v_settle */
    ora_conc cs[TB_NCONC_MAX];
    int nconc, conc_sel, hindered;
    double *ctot;
    double *tr1, *tr2;         /* transposed velocities (grid->G_dummy1/2) */
    /* open x faces (Inflow.c, Outflow.c): planes NY*NZ, [k][j] j fastest */
    tb_open_bc obc;
    double *u_in;              /* u->G_inflow */
    double *out[3], *out_old[3];   /* u, v, w G_outflow and G_outflow_old */
    double influx;             /* u->W_influx */
    double *ubc, *vbc, *wbc;   /* cell-centre velocities (G_data_bc) */

    ora_mat helm[4];
    double helm_dt[4];
    ora_mat pois;
    int pois_built;

    int nlev;
    ora_level lev[16];
    double *coarse_inv;        /* dense pseudo-inverse of the coarsest level */
    int coarse_n;
};

static inline size_t ora_idx(const ora_ctx *o, int i, int j, int k) {
    return ((size_t)k * (size_t)o->NY + (size_t)j) * (size_t)o->NX + (size_t)i;
}

/* entry (j, k) of an open-boundary plane */
static inline size_t ora_pidx(const ora_ctx *o, int j, int k) { return (size_t)k * (size_t)o->NY + (size_t)j; }

/* the inflow / outflow plane of quantity q (the selected field for c) when
 * that face is open, else NULL (ora_open.c) */
const double *ora_in_plane(const ora_ctx *o, int q);
const double *ora_out_plane(const ora_ctx *o, int q);
void ora_open_free(ora_ctx *o);
void ora_open_store_old(ora_ctx *o);
void ora_open_rk_average(ora_ctx *o, double old_frac, double new_frac);

/* grid / geometry (ora_grid.c) */
int ora_grid_init(ora_ctx *o);
int ora_status(const ora_ctx *o, int q, int i, int j, int k);
void ora_q_coords(const ora_ctx *o, int q, const double **xq, const double **yq,
                  const double **zq);

/* MyMath_interpolate_quantity, MyMath.c:13-77 */
double ora_interp2(double q0, double x0, double q1, double x1, double xi);
double ora_interp4(const double q[4], const double X1[4], const double X2[4],
                   double x1, double x2);

/* flow (ora_flow.c) and multigrid (ora_gmg.c) */
void ora_mat_free(ora_mat *m);
void ora_mat_apply(const ora_mat *m, const double *x, double *y);
int ora_gmg_build(ora_ctx *o);
void ora_gmg_free(ora_ctx *o);
void ora_vcycle(ora_ctx *o, const double *r_compact, double *z_compact);

#endif
