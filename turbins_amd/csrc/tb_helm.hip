// tb_helm.hip -- Helmholtz operators of the momentum and concentration
// predictors and their BiCGStab solve, all matrix-free on the device.
//
// Rows follow Velocity_u/v/w_form_LHS_matrix (Velocity.c:449-1471) and
// Conc_form_LHS_matrix (Conc.c:390-730) term by term, with the diagonal
// update of Velocity_modify_diagonal / Conc_modify_diagonal
// (Velocity.c:374-443, Conc.c:293-354) added to every non-IMMERSED row.
// Each row is summed in ascending column order, as PETSc's AIJ MatMult
// does, so one application is bit-identical to the assembled matrix's.
#include <cmath>

#include "tb_common.h"

struct HelmOp {
    Dims d;
    GridDev g;
    int q;
    const uint8_t *st;
    double Cx, Cy, Cz, ap_solid, mod;
    int wall[6];
    int particle;
    double robin_b_scale;   // 2/(Pe*dXi)
    double vs_abs;
};

__device__ inline int st_at(const Dims &d, const uint8_t *st, int i, int j, int k) {
    if (i < 0 || i > d.NX - 1 || j < 0 || j > d.NY - 1 || k < 0 || k > d.NZ - 1) return TB_OUTSIDE;
    return st[gidx(d, i, j, k)];
}

// y = A x for one non-IMMERSED row; returns false for IMMERSED rows.
__device__ inline bool helm_row_apply(const HelmOp &H, int i, int j, int k, long id, const double *__restrict__ x,
                                      double &y) {
    const Dims &d = H.d;
    const GridDev &g = H.g;
    int s = H.st[id];
    if (s == TB_IMMERSED) return false;
    bool row7 = (H.q == TB_Q_C) ? (s == TB_FLUID) : (s == TB_FLUID || s == TB_BUFFER_SOLID);
    if (!row7) {
        y = (H.ap_solid + H.mod) * x[id];
        return true;
    }
    const double Cx = H.Cx, Cy = H.Cy, Cz = H.Cz;
    double mpx, mpy, mpz, mex, mwx, mny, msy, mbz, mfz, ap;
    bool hb, hs, hw, he, hn, hf;
    if (H.q == TB_Q_U) {
        mpx = g.mxu[i]; mpy = g.myc[j]; mpz = g.mzc[k]; mex = g.mxc[i]; mwx = g.mxc[i - 1];
        mny = g.myv[j + 1]; msy = g.myv[j]; mbz = g.mzw[k]; mfz = g.mzw[k + 1];
        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mfz + mbz);
    } else if (H.q == TB_Q_V) {
        mpx = g.mxc[i]; mpy = g.myv[j]; mpz = g.mzc[k]; mex = g.mxu[i + 1]; mwx = g.mxu[i];
        mny = g.myc[j]; msy = g.myc[j - 1]; mbz = g.mzw[k]; mfz = g.mzw[k + 1];
        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
    } else if (H.q == TB_Q_W) {
        mpx = g.mxc[i]; mpy = g.myc[j]; mpz = g.mzw[k]; mex = g.mxu[i + 1]; mwx = g.mxu[i];
        mny = g.myv[j + 1]; msy = g.myv[j]; mbz = g.mzc[k - 1]; mfz = g.mzc[k];
        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
    } else {
        mpx = g.mxc[i]; mpy = g.myc[j]; mpz = g.mzc[k]; mex = g.mxu[i + 1]; mwx = g.mxu[i];
        mny = g.myv[j + 1]; msy = g.myv[j]; mbz = g.mzw[k]; mfz = g.mzw[k + 1];
        ap = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy) - Cz * mpz * (mbz + mfz);
    }
    double aw = Cx * mpx * mwx, ae = Cx * mpx * mex;
    double an = Cy * mpy * mny, as = Cy * mpy * msy;
    double ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
#define WALL(a, w) do { if (H.wall[w]) ap -= (a); else ap += (a); } while (0)
    int NX = d.NX, NY = d.NY, NZ = d.NZ;
    if (H.q == TB_Q_U) {
        hb = (k != 0); if (!hb) WALL(ab, 4);
        hs = (j != 0); if (!hs) WALL(as, 2);
        hw = st_at(d, H.st, i - 1, j, k) != TB_BOUNDARY;
        he = st_at(d, H.st, i + 1, j, k) != TB_BOUNDARY;
        hn = (j != NY - 2); if (!hn) WALL(an, 3);
        hf = (k != NZ - 2); if (!hf) WALL(af, 5);
    } else if (H.q == TB_Q_V) {
        hb = (k != 0); if (!hb) WALL(ab, 4);
        hs = st_at(d, H.st, i, j - 1, k) != TB_BOUNDARY;
        hw = (i != 0); if (!hw) WALL(aw, 0);
        he = (i < NX - 2); if (i == NX - 2) WALL(ae, 1);
        hn = st_at(d, H.st, i, j + 1, k) != TB_BOUNDARY;
        hf = (k != NZ - 2); if (!hf) WALL(af, 5);
    } else if (H.q == TB_Q_W) {
        hb = st_at(d, H.st, i, j, k - 1) != TB_BOUNDARY;
        hs = (j != 0); if (!hs) WALL(as, 2);
        hw = (i != 0); if (!hw) WALL(aw, 0);
        he = (i < NX - 2); if (i == NX - 2) WALL(ae, 1);
        hn = (j != NY - 2); if (!hn) WALL(an, 3);
        hf = st_at(d, H.st, i, j, k + 1) != TB_BOUNDARY;
    } else {
        hb = (k > 0); if (!hb) ap += ab;
        hs = (j > 0); if (!hs) ap += as;
        hw = (i > 0); if (!hw) ap += aw;
        he = (i < NX - 2); if (!he) ap += ae;
        hn = (j < NY - 2);
        if (!hn) {
            double coef = an;
            if (H.particle) {
                double b = H.robin_b_scale * mny;
                coef = an * (b - H.vs_abs) / (b + H.vs_abs);
            }
            ap += coef;
        }
        hf = (k < NZ - 2); if (!hf) ap += af;
    }
#undef WALL
    double apd = ap + H.mod;
    double r = 0.0;
    if (hb) r += ab * x[id - d.plane];
    if (hs) r += as * x[id - d.px];
    if (hw) r += aw * x[id - 1];
    r += apd * x[id];
    if (he) r += ae * x[id + 1];
    if (hn) r += an * x[id + d.px];
    if (hf) r += af * x[id + d.plane];
    y = r;
    return true;
}

__global__ void __launch_bounds__(256) k_helm_apply(HelmOp H, const double *__restrict__ x, double *__restrict__ y,
                                                    Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(H.d, i, j, k);
    double v;
    if (helm_row_apply(H, i, j, k, id, x, v)) y[id] = v;
}

// IMMERSED rows from the sorted IB row table
__global__ void k_ib_apply(const long *__restrict__ row, const int *__restrict__ cnt, const long *__restrict__ col,
                           const double *__restrict__ val, int n, const double *__restrict__ x,
                           double *__restrict__ y) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0;
    int m = cnt[a];
    for (int e = 0; e < m; e++) s += val[(long)a * 9 + e] * x[col[(long)a * 9 + e]];
    y[row[a]] = s;
}

static HelmOp make_op(tb_ctx *c, int q, double dt) {
    HelmOp H;
    H.d = c->d; H.g = c->g; H.q = q; H.st = c->st[q];
    double kappa = (q == TB_Q_C) ? c->cfg.pe : c->cfg.re;
    H.Cx = -1.0 / (kappa * c->dEta * c->dEta);
    H.Cy = -1.0 / (kappa * c->dXi * c->dXi);
    H.Cz = -1.0 / (kappa * c->dPhi * c->dPhi);
    H.ap_solid = -H.Cx * c->hgrid[15][0] - H.Cy * c->hgrid[16][0] - H.Cz * c->hgrid[17][0];
    H.mod = 1.0 / dt;
    for (int a = 0; a < 6; a++) H.wall[a] = c->cfg.wall_noslip[a];
    H.particle = (q == TB_Q_C && c->cfg.conc_type == TB_PARTICLE);
    H.robin_b_scale = 2.0 / (c->cfg.pe * c->dXi);
    H.vs_abs = fabs(c->cfg.v_settle);
    return H;
}

static int helm_apply(tb_ctx *c, const HelmOp &H, const double *x, double *y) {
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    hipLaunchKernelGGL(k_helm_apply, tiles_grid(t), tiles_block(), 0, c->s, H, x, y, t);
    int q = H.q;
    if (c->nib[q] > 0)
        hipLaunchKernelGGL(k_ib_apply, dim3((c->nib[q] + 255) / 256), dim3(256), 0, c->s, c->ib_row[q], c->ib_cnt[q],
                           c->ib_col[q], c->ib_val[q], c->nib[q], x, y);
    TB_LAUNCH_CHECK(c);
    return 0;
}

extern "C" int tb_helmholtz_apply(tb_ctx *c, int q, double dt, int f_in, int f_out) {
    if (!c || q < 0 || q > 3 || f_in < 0 || f_in >= TB_F_COUNT || f_out < 0 || f_out >= TB_F_COUNT || f_in == f_out) {
        if (c) c->err = "tb_helmholtz_apply: bad argument";
        return -1;
    }
    HelmOp H = make_op(c, q, dt);
    if (helm_apply(c, H, c->f[f_in], c->f[f_out])) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

// ---------------------------------------------------------------------------
// vector kernels with reductions (partials in c->red, RED_MAX_BLOCKS apart)

constexpr int VB = 1024;   // blocks of the grid-stride vector kernels
constexpr int VT = 256;

// r = b - y ; rhat = r ; p = v = 0 ; partials: r.r, b.b
__global__ void k_bicg_init(long n, const double *__restrict__ b, const double *__restrict__ y,
                            double *__restrict__ r, double *__restrict__ rh, double *__restrict__ p,
                            double *__restrict__ v, double *part) {
    __shared__ double sh[32];
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double bi = b[i];
        double ri = bi - y[i];
        r[i] = ri; rh[i] = ri; p[i] = 0.0; v[i] = 0.0;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = acc[0]; part[RED_MAX_BLOCKS + blockIdx.x] = acc[1]; }
}

// scalar updates on the device
__global__ void k_scalar(double *s, int op) {
    if (threadIdx.x != 0) return;
    switch (op) {
    case 0:   // start: rho = alpha = omega = 1 ; RHV holds rhat.r of the new residual
        s[S_RHO] = 1.0; s[S_ALPHA] = 1.0; s[S_OMEGA] = 1.0;
        break;
    case 1: { // beta = (rho1/rho)*(alpha/omega) ; rho = rho1 (rho1 in S_RHV)
        double rho1 = s[S_RHV];
        s[S_BETA] = (rho1 / s[S_RHO]) * (s[S_ALPHA] / s[S_OMEGA]);
        s[S_RHO] = rho1;
        break;
    }
    case 2:   // alpha = rho / (rhat.v) (in S_TMP0)
        s[S_ALPHA] = s[S_RHO] / s[S_TMP0];
        break;
    case 3:   // omega = (t.s)/(t.t)
        s[S_OMEGA] = (s[S_TT] > 0.0) ? s[S_TS] / s[S_TT] : 0.0;
        break;
    }
}

__global__ void k_bicg_p(long n, const double *__restrict__ r, const double *__restrict__ v, double *__restrict__ p,
                         const double *__restrict__ s) {
    double beta = s[S_BETA], omega = s[S_OMEGA];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

__global__ void k_bicg_s(long n, const double *__restrict__ r, const double *__restrict__ v, double *__restrict__ sv,
                         const double *__restrict__ s) {
    double alpha = s[S_ALPHA];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        sv[i] = r[i] - alpha * v[i];
}

// two dot products a.b and a.a (or c.d)
__global__ void k_dot2(long n, const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c2,
                       const double *__restrict__ d2, double *part) {
    __shared__ double sh[32];
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        acc[0] += a[i] * b[i];
        acc[1] += c2[i] * d2[i];
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = acc[0]; part[RED_MAX_BLOCKS + blockIdx.x] = acc[1]; }
}

// x += alpha p + omega s ; r = s - omega t ; partials r.r and rhat.r
__global__ void k_bicg_xr(long n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                          const double *__restrict__ sv, const double *__restrict__ t, const double *__restrict__ rh,
                          const double *__restrict__ s, double *part) {
    __shared__ double sh[32];
    double alpha = s[S_ALPHA], omega = s[S_OMEGA];
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double si = sv[i];
        x[i] += alpha * p[i] + omega * si;
        double ri = si - omega * t[i];
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += rh[i] * ri;
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = acc[0]; part[RED_MAX_BLOCKS + blockIdx.x] = acc[1]; }
}

__global__ void k_sub(long n, const double *__restrict__ b, const double *__restrict__ y, double *__restrict__ r,
                      double *part) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double ri = b[i] - y[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

static int read_scalars(tb_ctx *c, int first, int count) {
    TB_CHECK(c, hipMemcpyAsync(c->hscal + first, c->scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

// BiCGStab (van der Vorst 1992) without preconditioner, the method of
// KSPIBCGS + PCNONE (Solver.c:39-86, 269-322), nonzero initial guess = the
// current field, stopping at ||b - A x|| <= rtol ||b|| (PETSc's default
// test relative to the right-hand side).
extern "C" int tb_helmholtz_solve(tb_ctx *c, int q, double dt, tb_solve_stats *st) {
    if (!c || q < 0 || q > 3) { if (c) c->err = "tb_helmholtz_solve: bad quantity"; return -1; }
    HelmOp H = make_op(c, q, dt);
    long n = c->d.ntot;
    double *x = c->f[TB_F_U + q], *b = c->f[TB_F_RHS_U + q];
    double *r = c->kv[0], *rh = c->kv[1], *p = c->kv[2], *v = c->kv[3], *sv = c->kv[4], *t = c->kv[5];
    double rtol = c->cfg.helm_rtol;
    int maxit = c->cfg.helm_maxit;
    TB_CHECK(c, hipEventRecord(c->ev0, c->s));
    if (helm_apply(c, H, x, t)) return -1;
    hipLaunchKernelGGL(k_bicg_init, dim3(VB), dim3(VT), 0, c->s, n, b, t, r, rh, p, v, c->red);
    if (tb_reduce_finish(c, VB, 2, S_RR)) return -1;   // S_RR = r.r, S_PQ = b.b (next slot)
    hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 0);
    TB_LAUNCH_CHECK(c);
    if (read_scalars(c, S_RR, 2)) return -1;
    double rr = c->hscal[S_RR], bb = c->hscal[S_RR + 1];
    double bn = sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    double res0 = sqrt(rr) / bn, res = res0;
    // rho1 for the first iteration: rhat.r = r.r
    TB_CHECK(c, hipMemcpyAsync(c->scal + S_RHV, c->scal + S_RR, sizeof(double), hipMemcpyDeviceToDevice, c->s));
    int it = 0;
    bool conv = res <= rtol;
    while (!conv && it < maxit) {
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 1);
        hipLaunchKernelGGL(k_bicg_p, dim3(VB), dim3(VT), 0, c->s, n, r, v, p, c->scal);
        if (helm_apply(c, H, p, v)) return -1;
        hipLaunchKernelGGL(k_dot2, dim3(VB), dim3(VT), 0, c->s, n, rh, v, rh, v, c->red);
        if (tb_reduce_finish(c, VB, 1, S_TMP0)) return -1;
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 2);
        hipLaunchKernelGGL(k_bicg_s, dim3(VB), dim3(VT), 0, c->s, n, r, v, sv, c->scal);
        if (helm_apply(c, H, sv, t)) return -1;
        hipLaunchKernelGGL(k_dot2, dim3(VB), dim3(VT), 0, c->s, n, t, sv, t, t, c->red);
        if (tb_reduce_finish(c, VB, 2, S_TS)) return -1;   // S_TS, S_TT
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 3);
        hipLaunchKernelGGL(k_bicg_xr, dim3(VB), dim3(VT), 0, c->s, n, x, r, p, sv, t, rh, c->scal, c->red);
        if (tb_reduce_finish(c, VB, 2, S_RR)) return -1;   // S_RR = r.r ; S_PQ <- rhat.r
        TB_CHECK(c, hipMemcpyAsync(c->scal + S_RHV, c->scal + S_RR + 1, sizeof(double), hipMemcpyDeviceToDevice, c->s));
        TB_LAUNCH_CHECK(c);
        it++;
        if (read_scalars(c, S_RR, 1)) return -1;
        res = sqrt(c->hscal[S_RR]) / bn;
        conv = res <= rtol;
        if (!(res == res)) break;   // NaN guard
    }
    // true residual
    if (helm_apply(c, H, x, t)) return -1;
    hipLaunchKernelGGL(k_sub, dim3(VB), dim3(VT), 0, c->s, n, b, t, r, c->red);
    if (tb_reduce_finish(c, VB, 1, S_TMP1)) return -1;
    TB_CHECK(c, hipEventRecord(c->ev1, c->s));
    if (read_scalars(c, S_TMP1, 1)) return -1;
    double tres = sqrt(c->hscal[S_TMP1]) / bn;
    if (st) {
        st->iterations = it;
        st->converged = tres <= rtol * 10.0;
        st->fallback = 0;
        st->vcycles = 0;
        st->res0 = res0;
        st->res = tres;
        st->ms = tb_elapsed_ms(c);
    }
    return (tres <= rtol * 10.0) ? 0 : 1;
}
