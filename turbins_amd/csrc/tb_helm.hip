// tb_helm.hip -- Helmholtz operators of the momentum and concentration
// predictors and their BiCGStab solve, all matrix-free on the device.
//
// Rows follow Velocity_u/v/w_form_LHS_matrix (Velocity.c:449-1471) and
// Conc_form_LHS_matrix (Conc.c:390-730) term by term, with the diagonal
// update of Velocity_modify_diagonal / Conc_modify_diagonal
// (Velocity.c:374-443, Conc.c:293-354) added to every non-IMMERSED row.
// Each row is summed in ascending column order, as PETSc's AIJ MatMult
// does, so one application is bit-identical to the assembled matrix's.
//
// The operator kernel walks k inside each column tile: the x and y parts of
// every coefficient are formed once per column, the z parts once per plane
// (uniform across the block), and the k-1 / k / k+1 values stay in
// registers.  The neighbour-status tests of the reference reduce to index
// tests, because BOUNDARY nodes are exactly the box planes.
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

// wall rule applied to ap when a neighbour is absent:
// 0 nothing (normal component on the wall), 1 no-slip (ap -= a),
// 2 free-slip / Neumann (ap += a)
__device__ inline double wallmod(double ap, double a, int rule) {
    if (rule == 1) return ap - a;
    if (rule == 2) return ap + a;
    return ap;
}

// MODE 0: y = A x ; MODE 1: also partial w.y ; MODE 2: also partials x.y, y.y
// IMMERSED rows are produced by k_ib_apply, which runs first; here their y is
// only read for the dot products.
template <int MODE>
__global__ void __launch_bounds__(256) k_helm_km(HelmOp H, const double *__restrict__ x, double *__restrict__ y,
                                                 const double *__restrict__ w, double *part, KTiles t) {
    __shared__ double sh[32];
    double acc[2] = {0.0, 0.0};
    int i, j, k0, k1;
    bool act;
    ktile_column(t, i, j, k0, k1, act);
    if (act) {
        const Dims &d = H.d;
        const GridDev &g = H.g;
        const int NX = d.NX, NY = d.NY, NZ = d.NZ;
        const int q = H.q;
        const double Cx = H.Cx, Cy = H.Cy, Cz = H.Cz;
        const int im1 = i > 0 ? i - 1 : 0, ip1 = i + 1 < NX ? i + 1 : NX - 1;
        const int jm1 = j > 0 ? j - 1 : 0, jp1 = j + 1 < NY ? j + 1 : NY - 1;
        double mpx, mex, mwx, mpy, mny, msy;
        if (q == TB_Q_U) {
            mpx = g.mxu[i]; mex = g.mxc[i]; mwx = g.mxc[im1]; mpy = g.myc[j]; mny = g.myv[jp1]; msy = g.myv[j];
        } else if (q == TB_Q_V) {
            mpx = g.mxc[i]; mex = g.mxu[ip1]; mwx = g.mxu[i]; mpy = g.myv[j]; mny = g.myc[j]; msy = g.myc[jm1];
        } else {
            mpx = g.mxc[i]; mex = g.mxu[ip1]; mwx = g.mxu[i]; mpy = g.myc[j]; mny = g.myv[jp1]; msy = g.myv[j];
        }
        // ((-Cx mpx (mex+mwx)) - Cy mpy (mny+msy)) - Cz mpz (mbz+mfz), left to right
        const double axy = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy);
        const double aw = Cx * mpx * mwx, ae = Cx * mpx * mex;
        const double an = Cy * mpy * mny, as = Cy * mpy * msy;
        // per-column presence and wall rules of the x and y neighbours
        bool hs, hw, he, hn;
        int rs, rw, re, rn;
        const int *W = H.wall;
        if (q == TB_Q_U) {
            hs = j != 0;       rs = W[2] ? 1 : 2;
            hw = i - 1 >= 1;   rw = 0;
            he = i + 1 <= NX - 2; re = 0;
            hn = j != NY - 2;  rn = W[3] ? 1 : 2;
        } else if (q == TB_Q_V) {
            hs = j - 1 >= 1;   rs = 0;
            hw = i != 0;       rw = W[0] ? 1 : 2;
            he = i < NX - 2;   re = (i == NX - 2) ? (W[1] ? 1 : 2) : 0;
            hn = j + 1 <= NY - 2; rn = 0;
        } else if (q == TB_Q_W) {
            hs = j != 0;       rs = W[2] ? 1 : 2;
            hw = i != 0;       rw = W[0] ? 1 : 2;
            he = i < NX - 2;   re = (i == NX - 2) ? (W[1] ? 1 : 2) : 0;
            hn = j != NY - 2;  rn = W[3] ? 1 : 2;
        } else {
            hs = j > 0;        rs = 2;
            hw = i > 0;        rw = 2;
            he = i < NX - 2;   re = 2;
            hn = j < NY - 2;   rn = 2;
        }
        double cn = an;   // north wall term (Robin for settling particles, Conc.c:625-635)
        if (q == TB_Q_C && H.particle) {
            double b = H.robin_b_scale * mny;
            cn = an * (b - H.vs_abs) / (b + H.vs_abs);
        }
        const int rb = (q == TB_Q_W) ? 0 : (q == TB_Q_C ? 2 : (W[4] ? 1 : 2));
        const int rf = (q == TB_Q_W) ? 0 : (q == TB_Q_C ? 2 : (W[5] ? 1 : 2));
        const bool row_buffer = (q != TB_Q_C);   // BUFFER_SOLID rows carry the stencil for u, v, w
        const double diag_other = H.ap_solid + H.mod;

        // neighbour offsets clamped to the point itself at the box faces, so
        // every load is issued unconditionally and the loads overlap
        const long os = (j > 0) ? d.px : 0, on = (j + 1 < NY) ? d.px : 0;
        const long ow = (i > 0) ? 1 : 0, oe = (i + 1 < NX) ? 1 : 0;
        long id = gidx(d, i, j, k0);
        double xm = x[id - ((k0 > 0) ? d.plane : 0)];
        double xc = x[id];
        for (int k = k0; k < k1; k++) {
            double xp = x[id + ((k + 1 < NZ) ? d.plane : 0)];
            const double xs = x[id - os], xw = x[id - ow], xe = x[id + oe], xn = x[id + on];
            const double wv = (MODE == 1) ? w[id] : 0.0;
            int s = H.st[id];
            double yv;
            if (s == TB_IMMERSED) {
                yv = (MODE > 0) ? y[id] : 0.0;   // rare: only waves touching the IB layer wait here
            } else if (s == TB_FLUID || (row_buffer && s == TB_BUFFER_SOLID)) {
                const int km1 = k > 0 ? k - 1 : 0, kp1 = k + 1 < NZ ? k + 1 : NZ - 1;
                double mpz, mbz, mfz;
                if (q == TB_Q_W) { mpz = g.mzw[k]; mbz = g.mzc[km1]; mfz = g.mzc[k]; }
                else { mpz = g.mzc[k]; mbz = g.mzw[k]; mfz = g.mzw[kp1]; }
                double ap = axy - Cz * mpz * (mbz + mfz);
                const double ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
                bool hb, hf;
                if (q == TB_Q_W) { hb = k - 1 >= 1; hf = k + 1 <= NZ - 2; }
                else if (q == TB_Q_C) { hb = k > 0; hf = k < NZ - 2; }
                else { hb = k != 0; hf = k != NZ - 2; }
                // wall terms in the reference's order: back, south, west, east, north, front
                if (!hb) ap = wallmod(ap, ab, rb);
                if (!hs) ap = wallmod(ap, as, rs);
                if (!hw) ap = wallmod(ap, aw, rw);
                if (!he) ap = wallmod(ap, ae, re);
                if (!hn) ap = (q == TB_Q_C) ? ap + cn : wallmod(ap, an, rn);
                if (!hf) ap = wallmod(ap, af, rf);
                double r = 0.0;
                r = hb ? r + ab * xm : r;
                r = hs ? r + as * xs : r;
                r = hw ? r + aw * xw : r;
                r += (ap + H.mod) * xc;
                r = he ? r + ae * xe : r;
                r = hn ? r + an * xn : r;
                r = hf ? r + af * xp : r;
                y[id] = r;
                yv = r;
            } else {
                yv = diag_other * xc;
                y[id] = yv;
            }
            if (MODE == 1) acc[0] += wv * yv;
            if (MODE == 2) { acc[0] += xc * yv; acc[1] += yv * yv; }
            xm = xc;
            xc = xp;
            id += d.plane;
        }
    }
    if (MODE > 0) {
        block_sum<2>(acc, sh);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            part[blockIdx.x] = acc[0];
            part[RED_MAX_BLOCKS + blockIdx.x] = acc[1];
        }
    }
}

// IMMERSED rows from the sorted IB row table.  With target > 0 each row is
// multiplied by target / (its diagonal): inside the Krylov solve this
// rescales the IB equations (right-hand side 0) to the size of the fluid
// rows, which takes BiCGStab from ~15-20 to 3-4 iterations; any scaled
// residual >= the true one, so the stopping test stays conservative.
__global__ void k_ib_apply(const long *__restrict__ row, const int *__restrict__ cnt, const long *__restrict__ col,
                           const double *__restrict__ val, int n, const double *__restrict__ x,
                           double *__restrict__ y, double target) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0, dg = 1.0;
    long r = row[a];
    int m = cnt[a];
    for (int e = 0; e < m; e++) {
        long cidx = col[(long)a * 9 + e];
        double v = val[(long)a * 9 + e];
        if (cidx == r) dg = v;
        s += v * x[cidx];
    }
    double f = (target > 0.0) ? target / dg : 1.0;
    y[r] = (f > 1.0) ? f * s : s;
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

// y = A x (scaled: IB rows rescaled for the Krylov solve); mode 1/2 leave
// partial sums in c->red.  Returns the number of partials, or -1.
static int helm_apply(tb_ctx *c, const HelmOp &H, const double *x, double *y, bool scaled, int mode,
                      const double *w = nullptr) {
    int q = H.q;
    if (c->nib[q] > 0)
        hipLaunchKernelGGL(k_ib_apply, dim3((c->nib[q] + 255) / 256), dim3(256), 0, c->s, c->ib_row[q], c->ib_cnt[q],
                           c->ib_col[q], c->ib_val[q], c->nib[q], x, y, scaled ? H.mod : 0.0);
    KTiles t = make_ktiles(c->d.NX, c->d.NY, c->d.NZ);
    if (c->prof_on) tb_prof_begin(c, TB_K_HELM_OP);
    if (mode == 0)
        hipLaunchKernelGGL(k_helm_km<0>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t);
    else if (mode == 1)
        hipLaunchKernelGGL(k_helm_km<1>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t);
    else
        hipLaunchKernelGGL(k_helm_km<2>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t);
    if (c->prof_on) tb_prof_end(c, TB_K_HELM_OP);
    TB_LAUNCH_CHECK(c);
    return (int)(t.per_xcd * NXCD);
}

extern "C" int tb_helmholtz_apply(tb_ctx *c, int q, double dt, int f_in, int f_out) {
    if (!c || q < 0 || q > 3 || f_in < 0 || f_in >= TB_F_COUNT || f_out < 0 || f_out >= TB_F_COUNT || f_in == f_out) {
        if (c) c->err = "tb_helmholtz_apply: bad argument";
        return -1;
    }
    HelmOp H = make_op(c, q, dt);
    if (helm_apply(c, H, c->f[f_in], c->f[f_out], false, 0) < 0) return -1;
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
    case 0:   // start: rho = alpha = omega = 1
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

__global__ void k_copy_slot(double *s, int dst, int src) {
    if (threadIdx.x == 0) s[dst] = s[src];
}

static int read_scalars(tb_ctx *c, int first, int count) {
    TB_CHECK(c, hipMemcpyAsync(c->hscal + first, c->scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

// BiCGStab (van der Vorst 1992), the method of KSPIBCGS (Solver.c:39-86,
// 269-322), with the IB equations rescaled (see k_ib_apply); nonzero initial
// guess = the current field; stops at ||b - A x|| <= rtol ||b|| (PETSc's
// default test relative to the right-hand side), checked on the true
// residual at the end.
extern "C" int tb_helmholtz_solve(tb_ctx *c, int q, double dt, tb_solve_stats *st) {
    if (!c || q < 0 || q > 3) { if (c) c->err = "tb_helmholtz_solve: bad quantity"; return -1; }
    HelmOp H = make_op(c, q, dt);
    long n = c->d.ntot;
    double *x = c->f[TB_F_U + q], *b = c->f[TB_F_RHS_U + q];
    double *r = c->kv[0], *rh = c->kv[1], *p = c->kv[2], *v = c->kv[3], *sv = c->kv[4], *t = c->kv[5];
    double rtol = c->cfg.helm_rtol;
    int maxit = c->cfg.helm_maxit;
    TB_CHECK(c, hipEventRecord(c->ev0, c->s));
    if (helm_apply(c, H, x, t, true, 0) < 0) return -1;
    hipLaunchKernelGGL(k_bicg_init, dim3(VB), dim3(VT), 0, c->s, n, b, t, r, rh, p, v, c->red);
    if (tb_reduce_finish(c, VB, 2, S_RR)) return -1;   // S_RR = r.r, S_RR+1 = b.b
    hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 0);
    TB_LAUNCH_CHECK(c);
    if (read_scalars(c, S_RR, 2)) return -1;
    double rr = c->hscal[S_RR], bb = c->hscal[S_RR + 1];
    double bn = sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    double res0 = sqrt(rr) / bn, res = res0;
    hipLaunchKernelGGL(k_copy_slot, dim3(1), dim3(64), 0, c->s, c->scal, S_RHV, S_RR);   // rhat.r = r.r
    int it = 0;
    bool conv = res <= rtol;
    while (!conv && it < maxit) {
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 1);
        hipLaunchKernelGGL(k_bicg_p, dim3(VB), dim3(VT), 0, c->s, n, r, v, p, c->scal);
        int nb = helm_apply(c, H, p, v, true, 1, rh);                  // v = A p, rhat.v
        if (nb < 0 || tb_reduce_finish(c, nb, 1, S_TMP0)) return -1;
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 2);
        hipLaunchKernelGGL(k_bicg_s, dim3(VB), dim3(VT), 0, c->s, n, r, v, sv, c->scal);
        nb = helm_apply(c, H, sv, t, true, 2);                         // t = A s, t.s, t.t
        if (nb < 0 || tb_reduce_finish(c, nb, 2, S_TS)) return -1;
        hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, c->s, c->scal, 3);
        TB_PROF(c, TB_K_BICG_UPD,
                hipLaunchKernelGGL(k_bicg_xr, dim3(VB), dim3(VT), 0, c->s, n, x, r, p, sv, t, rh, c->scal, c->red));
        if (tb_reduce_finish(c, VB, 2, S_TMP1)) return -1;              // r.r, rhat.r
        hipLaunchKernelGGL(k_copy_slot, dim3(1), dim3(64), 0, c->s, c->scal, S_RHV, S_TMP2);
        TB_LAUNCH_CHECK(c);
        it++;
        if (read_scalars(c, S_TMP1, 1)) return -1;
        res = sqrt(c->hscal[S_TMP1]) / bn;
        conv = res <= rtol;
        if (!(res == res)) break;   // NaN guard
    }
    // true residual of the unscaled operator
    if (helm_apply(c, H, x, t, false, 0) < 0) return -1;
    hipLaunchKernelGGL(k_sub, dim3(VB), dim3(VT), 0, c->s, n, b, t, r, c->red);
    if (tb_reduce_finish(c, VB, 1, S_TMP1)) return -1;
    TB_CHECK(c, hipEventRecord(c->ev1, c->s));
    if (read_scalars(c, S_TMP1, 1)) return -1;
    double tres = sqrt(c->hscal[S_TMP1]) / bn;
    if (st) {
        st->iterations = it;
        st->converged = tres <= rtol;
        st->fallback = 0;
        st->vcycles = 0;
        st->res0 = res0;
        st->res = tres;
        st->ms = tb_elapsed_ms(c);
    }
    return (tres <= rtol) ? 0 : 1;
}
