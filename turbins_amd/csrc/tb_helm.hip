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
    // open x faces drop their wall rule: c at x = 0 with inflow, v, w and c
    // at x = Lx with outflow (Velocity.c:989, 1336; Conc.c:582-607)
    int open_w, open_e;
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
// MODE 3: residual only: partials |w - A x|^2 and |w|^2 with w = b, nothing
// stored except that IMMERSED rows must hold (A x) from k_ib_apply in y.
// MODE 4: MODE 3 and the increment dn = x - x0 over the box (the Gauss-Seidel
// solve's change to the field it started from).
// IMMERSED rows are produced by k_ib_apply, which runs first; here their y is
// only read for the dot products.
template <int MODE>
__global__ void __launch_bounds__(256) k_helm_km(HelmOp H, const double *__restrict__ x, double *__restrict__ y,
                                                 const double *__restrict__ w, double *part, KTiles t,
                                                 const double *__restrict__ x0, float *__restrict__ dn) {
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
            he = i < NX - 2;   re = (i == NX - 2) ? (H.open_e ? 0 : (W[1] ? 1 : 2)) : 0;
            hn = j + 1 <= NY - 2; rn = 0;
        } else if (q == TB_Q_W) {
            hs = j != 0;       rs = W[2] ? 1 : 2;
            hw = i != 0;       rw = W[0] ? 1 : 2;
            he = i < NX - 2;   re = (i == NX - 2) ? (H.open_e ? 0 : (W[1] ? 1 : 2)) : 0;
            hn = j != NY - 2;  rn = W[3] ? 1 : 2;
        } else {
            hs = j > 0;        rs = 2;
            hw = i > 0;        rw = H.open_w ? 0 : 2;
            he = i < NX - 2;   re = H.open_e ? 0 : 2;
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
            const double wv = (MODE == 1 || MODE >= 3) ? __builtin_nontemporal_load(&w[id]) : 0.0;
            const double x0v = (MODE == 4) ? __builtin_nontemporal_load(&x0[id]) : 0.0;
            const int s = H.st[id];
            // the fluid row is formed for every node and selected afterwards:
            // no status-dependent branch sits between the loads and the math
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
            const bool fluid = (s == TB_FLUID) || (row_buffer && s == TB_BUFFER_SOLID);
            double yv = fluid ? r : diag_other * xc;
            if (s == TB_IMMERSED) {
                yv = (MODE > 0) ? y[id] : 0.0;   // row written by k_ib_apply
            } else if (MODE < 3) {
                st_stream(&y[id], yv);
            }
            if (MODE == 4) st_stream(&dn[id], (float)(xc - x0v));
            if (MODE == 1) acc[0] += wv * yv;
            if (MODE >= 3) { acc[0] += (wv - yv) * (wv - yv); acc[1] += wv * wv; }
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
    const long r = row[a];
    const int m = cnt[a];
    // loads first, then the sum in column order (as in k_ib_gs)
    long ci[9];
    double cv[9], xv[9];
#pragma unroll
    for (int e = 0; e < 9; e++) {
        ci[e] = e < m ? col[(long)a * 9 + e] : r;
        cv[e] = e < m ? val[(long)a * 9 + e] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 9; e++) xv[e] = e < m ? x[ci[e]] : 0.0;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        if (e >= m) continue;
        if (ci[e] == r) dg = cv[e];
        s += cv[e] * xv[e];
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
    H.open_w = (q == TB_Q_C) && c->obc.inflow;
    H.open_e = (q != TB_Q_U) && c->obc.outflow;
    return H;
}

// y = A x (scaled: IB rows rescaled for the Krylov solve); mode 1/2/3 leave
// partial sums in c->red.  Returns the number of partials, or -1.
static int helm_apply(tb_ctx *c, const HelmOp &H, const double *x, double *y, bool scaled, int mode,
                      const double *w = nullptr, const double *x0 = nullptr, float *dn = nullptr) {
    int q = H.q;
    if (c->nib[q] > 0)
        hipLaunchKernelGGL(k_ib_apply, dim3((c->nib[q] + 255) / 256), dim3(256), 0, c->s, c->ib_row[q], c->ib_cnt[q],
                           c->ib_col[q], c->ib_val[q], c->nib[q], x, y, scaled ? H.mod : 0.0);
    KTiles t = make_ktiles(c->d.NX, c->d.NY, c->d.NZ);
    if (c->prof_on) tb_prof_begin(c, TB_K_HELM_OP);
    if (mode == 0)
        hipLaunchKernelGGL(k_helm_km<0>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t, x0, dn);
    else if (mode == 1)
        hipLaunchKernelGGL(k_helm_km<1>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t, x0, dn);
    else if (mode == 2)
        hipLaunchKernelGGL(k_helm_km<2>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t, x0, dn);
    else if (mode == 3)
        hipLaunchKernelGGL(k_helm_km<3>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t, x0, dn);
    else
        hipLaunchKernelGGL(k_helm_km<4>, ktiles_grid(t), tiles_block(), 0, c->s, H, x, y, w, c->red, t, x0, dn);
    if (c->prof_on) tb_prof_end(c, TB_K_HELM_OP);
    TB_LAUNCH_CHECK(c);
    return (int)(t.per_xcd * NXCD);
}

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel solve of the Helmholtz system.  The rows are
// dominated by 1/dt (at 512x256x128, Re 2000, dt 1e-3 the off-diagonal sum
// of a fluid row is ~8% of its diagonal), so one red-black sweep reduces the
// error by about two orders of magnitude and a sweep moves ~33 bytes per
// point against ~154 for a BiCGStab iteration.  The solve stops on the same
// test as the reference's KSP, ||b - A x|| <= rtol ||b|| with the true
// residual of the unscaled operator, and hands over to BiCGStab if the
// sweeps stop contracting.

// Per-axis parts of a row, with the terms and wall rules of k_helm_km
// (Velocity.c:449-1471, Conc.c:390-730): d is the axis's share of the
// diagonal with its wall rules folded in, am / ap the minus / plus neighbour
// coefficients (zero where the neighbour is absent).  The diagonal of the
// Gauss-Seidel update is (d_x + d_y) + d_z + 1/dt; it differs from the
// assembled row's only in rounding (the rows sum the same terms in another
// order), so a sweep's fixed point solves the assembled system to ~1e-16.
struct AxisGS {
    double d, am, ap;
};

__device__ inline AxisGS axis_gs(double d, double am, bool hm, int rm, double ap, double wp, bool hp, int rp) {
    AxisGS a;
    if (!hm) d = wallmod(d, am, rm);
    if (!hp) d = wallmod(d, wp, rp);
    a.d = d;
    a.am = hm ? am : 0.0;
    a.ap = hp ? ap : 0.0;
    return a;
}

__device__ inline AxisGS axis_x(const HelmOp &H, int i) {
    const GridDev &g = H.g;
    const int NX = H.d.NX, q = H.q;
    const int *W = H.wall;
    const int im1 = i > 0 ? i - 1 : 0, ip1 = i + 1 < NX ? i + 1 : NX - 1;
    double mpx, mex, mwx;
    if (q == TB_Q_U) { mpx = g.mxu[i]; mex = g.mxc[i]; mwx = g.mxc[im1]; }
    else { mpx = g.mxc[i]; mex = g.mxu[ip1]; mwx = g.mxu[i]; }
    const double Cx = H.Cx;
    const double t = -Cx * mpx * (mex + mwx), aw = Cx * mpx * mwx, ae = Cx * mpx * mex;
    if (q == TB_Q_U) return axis_gs(t, aw, i - 1 >= 1, 0, ae, ae, i + 1 <= NX - 2, 0);
    if (q == TB_Q_C) return axis_gs(t, aw, i > 0, H.open_w ? 0 : 2, ae, ae, i < NX - 2, H.open_e ? 0 : 2);
    return axis_gs(t, aw, i != 0, W[0] ? 1 : 2, ae, ae, i < NX - 2,
                   (i == NX - 2) ? (H.open_e ? 0 : (W[1] ? 1 : 2)) : 0);
}

__device__ inline AxisGS axis_y(const HelmOp &H, int j) {
    const GridDev &g = H.g;
    const int NY = H.d.NY, q = H.q;
    const int *W = H.wall;
    const int jm1 = j > 0 ? j - 1 : 0, jp1 = j + 1 < NY ? j + 1 : NY - 1;
    double mpy, mny, msy;
    if (q == TB_Q_V) { mpy = g.myv[j]; mny = g.myc[j]; msy = g.myc[jm1]; }
    else { mpy = g.myc[j]; mny = g.myv[jp1]; msy = g.myv[j]; }
    const double Cy = H.Cy;
    const double t = -(Cy * mpy * (mny + msy)), as = Cy * mpy * msy, an = Cy * mpy * mny;
    double wn = an;   // top-wall coefficient: the Robin form for settling particles (Conc.c:560-585)
    if (q == TB_Q_C && H.particle) {
        const double bb = H.robin_b_scale * mny;
        wn = an * (bb - H.vs_abs) / (bb + H.vs_abs);
    }
    if (q == TB_Q_V) return axis_gs(t, as, j - 1 >= 1, 0, an, wn, j + 1 <= NY - 2, 0);
    if (q == TB_Q_C) return axis_gs(t, as, j > 0, 2, an, wn, j < NY - 2, 2);
    return axis_gs(t, as, j != 0, W[2] ? 1 : 2, an, wn, j != NY - 2, W[3] ? 1 : 2);
}

__device__ inline AxisGS axis_z(const HelmOp &H, int k) {
    const GridDev &g = H.g;
    const int NZ = H.d.NZ, q = H.q;
    const int *W = H.wall;
    const int km1 = k > 0 ? k - 1 : 0, kp1 = k + 1 < NZ ? k + 1 : NZ - 1;
    double mpz, mbz, mfz;
    if (q == TB_Q_W) { mpz = g.mzw[k]; mbz = g.mzc[km1]; mfz = g.mzc[k]; }
    else { mpz = g.mzc[k]; mbz = g.mzw[k]; mfz = g.mzw[kp1]; }
    const double Cz = H.Cz;
    const double t = -(Cz * mpz * (mbz + mfz)), ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
    if (q == TB_Q_W) return axis_gs(t, ab, k - 1 >= 1, 0, af, af, k + 1 <= NZ - 2, 0);
    if (q == TB_Q_C) return axis_gs(t, ab, k > 0, 2, af, af, k < NZ - 2, 2);
    return axis_gs(t, ab, k != 0, W[4] ? 1 : 2, af, af, k != NZ - 2, W[5] ? 1 : 2);
}

// Gauss-Seidel value of a row of status s from its own current value xp and
// its six neighbour values (clamped to the point itself at the box faces,
// where the coefficient is zero): IMMERSED rows keep xp, other non-fluid
// rows take b / (ap_solid + 1/dt)
__device__ inline double gs_value(const HelmOp &H, const AxisGS &X, const AxisGS &Y, const AxisGS &Z, int s,
                                  double bi, double xp, double xb, double xs, double xw, double xe, double xn,
                                  double xf) {
    const bool fluid = (s == TB_FLUID) || (H.q != TB_Q_C && s == TB_BUFFER_SOLID);
    double o = Z.am * xb;
    o = fma(Y.am, xs, o);
    o = fma(X.am, xw, o);
    o = fma(X.ap, xe, o);
    o = fma(Y.ap, xn, o);
    o = fma(Z.ap, xf, o);
    const double num = fluid ? bi - o : bi;
    const double den = fluid ? ((X.d + Y.d) + Z.d) + H.mod : H.ap_solid + H.mod;
    const double v = num / den;
    return s == TB_IMMERSED ? xp : v;
}

// One red-black sweep pair over the non-IMMERSED rows in a single pass:
// xn = black(red(xo)), red = (i + j + k) even, updated first.  Every red
// value reads only black neighbours of xo and every black value only red
// neighbours of the red pass (a neighbour clamped to the point itself reads
// the point's own current value), so the pair is the same iteration as two
// in-place colour passes (equal to rounding: the diagonals are summed per
// axis, see AxisGS), while xo and b are read from HBM once and xn written
// once (the in-place form streams both arrays twice).
//
// A 512-thread block owns a 64 x 16 column tile and a chunk of GS_TK planes
// and walks k upwards through two LDS rings: four planes of xo with a
// two-point halo, and three of the red-pass values (black positions holding
// xo) with a one-point halo.  Per plane m, with two barriers: store plane m-2
// (both of its passes are done) and red pass on plane m (tile + halo,
// recomputed by the neighbouring tiles); black pass on plane m-1 (tile).
// The fourth xo plane lets the next plane's staging overlap the black pass.
// The chunk recomputes the red pass on one plane below and above.  The
// xo plane, statuses and right-hand sides of the next plane are fetched into
// registers one step ahead, so their latency overlaps the passes; row
// coefficients come from per-axis tables built in LDS once per block.
// 32-plane chunks: at 512 x 256 x 128 the 8 x 16 x 4 tiles are exactly the
// blocks resident at once (two per CU at 75 KB of LDS), so no second round
// runs on a partly idle chip, and the chunk ends recompute 1/16 of the
// planes instead of 1/8 (126 -> 115 us per pair against 16-plane chunks).
constexpr int GS_TX = 64, GS_TY = 16, GS_TK = 32;
constexpr int GS_OX = GS_TX + 4, GS_OY = GS_TY + 4;   // xo ring: halo 2
constexpr int GS_HX = GS_TX + 2, GS_HY = GS_TY + 2;   // red-pass ring: halo 1
constexpr int GS_OZ = GS_TK + 4;                      // z table: k0-2 .. k0+GS_TK+1
constexpr int GS_ORING = 4;                           // xo planes staged: m-1, m, m+1 and the next
constexpr int GS_THREADS = 512;
constexpr int GS_OLD_PER = (GS_OY * GS_OX + GS_THREADS - 1) / GS_THREADS;
constexpr int GS_RED = GS_HY * (GS_HX / 2);           // red points of one halo plane
constexpr int GS_RED_PER = (GS_RED + GS_THREADS - 1) / GS_THREADS;
static_assert(GS_TX * GS_TY == 2 * GS_THREADS, "one black point per thread");
static_assert(GS_TX % 2 == 0, "tile origins must be even");

struct GsTiles {
    int tx, ty, kc;
    int n, per_xcd;
};

// The tiles cover the box without its last plane along each axis (see
// k_gs_last_planes), so 512 x 256 x 128 cells take exactly 8 x 16 x 4 tiles
// instead of 9 x 17 x 9 with nearly empty ones walking the full plane chain.
static GsTiles make_gs_tiles(const Dims &d) {
    GsTiles t;
    t.tx = (d.NX - 1 + GS_TX - 1) / GS_TX;
    t.ty = (d.NY - 1 + GS_TY - 1) / GS_TY;
    t.kc = (d.NZ - 1 + GS_TK - 1) / GS_TK;
    t.n = t.tx * t.ty * t.kc;
    t.per_xcd = (t.n + NXCD - 1) / NXCD;
    return t;
}

// Neighbours are read without clamping at the box faces: every row whose
// neighbour index leaves the box either has a zero coefficient there or is
// not a FLUID row (the last plane of each axis is BOUNDARY), and the LDS
// positions outside the box hold zeros or finite values of other planes.
// Element offsets are 32-bit (the host checks the field size).  ADD: the
// pair starts from xo + dd (the field plus the last solve's increment, an
// extrapolated initial guess) without forming it in memory.
template <bool ADD>
__global__ void __launch_bounds__(GS_THREADS) k_helm_gs_pair(HelmOp H, const double *__restrict__ xo,
                                                             double *__restrict__ xn, const double *__restrict__ b,
                                                             GsTiles T, const float *__restrict__ dd) {
    __shared__ double so[GS_ORING][GS_OY * GS_OX];
    __shared__ double sr[3][GS_HY * GS_HX];
    __shared__ AxisGS cx[GS_OX], cy[GS_OY], cz[GS_OZ];
    const Dims &d = H.d;
    const int NX = d.NX, NY = d.NY, NZ = d.NZ, px = d.px;
    const int plane = (int)d.plane;
    const int bid = blockIdx.x;
    const int lt = (bid % NXCD) * T.per_xcd + bid / NXCD;   // consecutive tiles share an XCD's L2
    if (lt >= T.n) return;                                  // whole block
    const int tx = lt % T.tx, rest = lt / T.tx;
    const int i0 = tx * GS_TX, j0 = (rest % T.ty) * GS_TY;
    const int k0 = (rest / T.ty) * GS_TK;
    const int k1 = k0 + GS_TK < NZ - 1 ? k0 + GS_TK : NZ - 1;   // the last plane: k_gs_last_planes
    const int t = threadIdx.x;
    const uint8_t *__restrict__ st = H.st;

    for (int e = t; e < GS_ORING * GS_OY * GS_OX; e += GS_THREADS) (&so[0][0])[e] = 0.0;
    for (int e = t; e < 3 * GS_HY * GS_HX; e += GS_THREADS) (&sr[0][0])[e] = 0.0;
    for (int e = t; e < GS_OX + GS_OY + GS_OZ; e += GS_THREADS) {
        if (e < GS_OX) {
            const int i = i0 - 2 + e;
            if (i >= 0 && i < NX) cx[e] = axis_x(H, i);
        } else if (e < GS_OX + GS_OY) {
            const int j = j0 - 2 + (e - GS_OX);
            if (j >= 0 && j < NY) cy[e - GS_OX] = axis_y(H, j);
        } else {
            const int k = k0 - 2 + (e - GS_OX - GS_OY);
            if (k >= 0 && k < NZ) cz[e - GS_OX - GS_OY] = axis_z(H, k);
        }
    }
    __syncthreads();   // zeros in place before the first planes are staged

    // xo plane staging: in-plane offset of each of this thread's halo-2
    // points (-1 outside the box); the LDS index is e itself
    int ooff[GS_OLD_PER];
#pragma unroll
    for (int u = 0; u < GS_OLD_PER; u++) {
        const int e = t + u * GS_THREADS;
        const int r = e / GS_OX, cc = e - r * GS_OX;
        const int i = i0 - 2 + cc, j = j0 - 2 + r;
        ooff[u] = (e < GS_OY * GS_OX && i >= 0 && i < NX && j >= 0 && j < NY) ? j * px + i : -1;
    }
    // the loads of a plane stay in flight until put_old (with ADD the sum is
    // formed there too, not between the loads)
    double ro[GS_OLD_PER];
    float rd[GS_OLD_PER];   // widened where the sum is formed, so the load is not waited for here
    auto fetch_old = [&](int m) {
#pragma unroll
        for (int u = 0; u < GS_OLD_PER; u++) {
            ro[u] = ooff[u] >= 0 ? xo[m * plane + ooff[u]] : 0.0;
            if (ADD) rd[u] = ooff[u] >= 0 ? dd[m * plane + ooff[u]] : 0.0f;
        }
    };
    auto put_old = [&](int m) {
        double *S = so[m % GS_ORING];
#pragma unroll
        for (int u = 0; u < GS_OLD_PER; u++) {
            const int e = t + u * GS_THREADS;
            if (e < GS_OY * GS_OX) S[e] = ADD ? ro[u] + (double)rd[u] : ro[u];
        }
    };
    // this thread's red points of the halo-1 plane: row, column pair, y
    int rrow[GS_RED_PER], rc2[GS_RED_PER];
#pragma unroll
    for (int u = 0; u < GS_RED_PER; u++) {
        const int e = t + u * GS_THREADS;
        rrow[u] = e < GS_RED ? e / (GS_HX / 2) : 0;
        rc2[u] = e < GS_RED ? e - rrow[u] * (GS_HX / 2) : -1;
    }
    // status (-1 outside the box) and b of the red points of plane m
    int rs[GS_RED_PER];
    double rb[GS_RED_PER];
    auto fetch_red = [&](int m, int *s_, double *b_) {
#pragma unroll
        for (int u = 0; u < GS_RED_PER; u++) {
            const int j = j0 - 1 + rrow[u];
            const int i = i0 - 1 + 2 * rc2[u] + ((1 + j + m) & 1);   // i + j + m even (i0 is even)
            s_[u] = -1;
            b_[u] = 0.0;
            if (rc2[u] >= 0 && i >= 0 && i < NX && j >= 0 && j < NY) {
                const int id = m * plane + j * px + i;
                s_[u] = st[id];
                b_[u] = b[id];
            }
        }
    };
#pragma unroll
    for (int u = 0; u < GS_RED_PER; u++) { rs[u] = -1; rb[u] = 0.0; }
    // this thread's black point of plane k
    const int br = t / (GS_TX / 2), bc2 = t % (GS_TX / 2);
    const int bj = j0 + br;
    int bs = -1;
    double bbv = 0.0;
    auto fetch_black = [&](int k, int &s_, double &b_) {
        const int i = i0 + 2 * bc2 + ((1 + bj + k) & 1);
        s_ = -1;
        b_ = 0.0;
        if (i < NX && bj < NY) {
            const int id = k * plane + bj * px + i;
            s_ = st[id];
            b_ = b[id];
        }
    };

    // both colours of plane k, from its sr slot to xn (coalesced rows)
    auto write_plane = [&](int k) {
        if (k < k0 || k >= k1) return;
        const double *R = sr[k % 3];
        for (int e = t; e < GS_TY * GS_TX; e += GS_THREADS) {
            const int r = e / GS_TX, cc = e - r * GS_TX;
            const int i = i0 + cc, j = j0 + r;
            if (i < NX && j < NY) st_stream(&xn[k * plane + j * px + i], R[(r + 1) * GS_HX + cc + 1]);
        }
    };

    if (k0 >= 2) { fetch_old(k0 - 2); put_old(k0 - 2); }
    if (k0 >= 1) { fetch_old(k0 - 1); put_old(k0 - 1); }
    fetch_old(k0);
    if (k0 >= 1) fetch_red(k0 - 1, rs, rb);
    for (int m = k0 - 1; m <= k1; m++) {
        if (m + 1 < NZ) put_old(m + 1);
        if (m + 2 < NZ) fetch_old(m + 2);
        int rs_n[GS_RED_PER];
        double rb_n[GS_RED_PER];
#pragma unroll
        for (int u = 0; u < GS_RED_PER; u++) { rs_n[u] = -1; rb_n[u] = 0.0; }
        if (m + 1 < NZ) fetch_red(m + 1, rs_n, rb_n);
        int bs_n = -1;
        double bb_n = 0.0;
        if (m >= k0 && m < k1) fetch_black(m, bs_n, bb_n);
        __syncthreads();
        // plane m - 2 is final (its black pass ran in the previous
        // iteration) and its sr slot is next written by the red pass of m + 1
        write_plane(m - 2);
        if (m >= 0 && m < NZ) {
            // red pass on plane m over the tile + halo 1
            const double *S = so[m % GS_ORING], *Sb = so[(m + GS_ORING - 1) % GS_ORING],
                         *Sf = so[(m + 1) % GS_ORING];
            double *R = sr[m % 3];
            const AxisGS Z = cz[m - k0 + 2];
#pragma unroll
            for (int u = 0; u < GS_RED_PER; u++) {
                const int s = rs[u];
                if (s < 0) continue;
                const int r = rrow[u];
                const int cc = 2 * rc2[u] + ((j0 + r + m) & 1);
                const int o = (r + 1) * GS_OX + cc + 1;
                R[r * GS_HX + cc] = gs_value(H, cx[cc + 1], cy[r + 1], Z, s, rb[u], S[o], Sb[o], S[o - GS_OX],
                                             S[o - 1], S[o + 1], S[o + GS_OX], Sf[o]);
            }
        }
        __syncthreads();
        const int k = m - 1;
        if (k >= k0 && k < k1 && bs >= 0) {
            // black pass on plane k over the tile, into the black positions of sr
            double *R = sr[k % 3];
            const double *Rb = sr[(k + 2) % 3], *Rf = sr[(k + 1) % 3];
            const int cc = 2 * bc2 + ((1 + bj + k) & 1);
            const int h = (br + 1) * GS_HX + cc + 1;
            R[h] = gs_value(H, cx[cc + 2], cy[br + 2], cz[k - k0 + 2], bs, bbv,
                            so[k % GS_ORING][(br + 2) * GS_OX + cc + 2], Rb[h], R[h - GS_HX], R[h - 1], R[h + 1],
                            R[h + GS_HX], Rf[h]);
        }
        // no barrier here: the next iteration's staging writes the xo slot of
        // plane m + 2, which no pass of this iteration reads (GS_ORING = 4),
        // and its first barrier precedes every read of this black pass
#pragma unroll
        for (int u = 0; u < GS_RED_PER; u++) { rs[u] = rs_n[u]; rb[u] = rb_n[u]; }
        bs = bs_n;
        bbv = bb_n;
    }
    __syncthreads();
    write_plane(k1 - 1);
}

// The last plane of each axis, outside the sweep tiles.  Its points are
// BOUNDARY by classification: their rows read ap_solid q = b and no other
// row couples to them (zero coefficients), so their Gauss-Seidel value
// b / (ap_solid + 1/dt), the expression of gs_value, does not depend on x and
// is written into both iterate buffers once per gs_pairs call.  Points where
// two planes meet are written twice with the same value.
__global__ void k_gs_last_planes(HelmOp H, const double *__restrict__ b, double *__restrict__ x1,
                                 double *__restrict__ x2) {
    const Dims &d = H.d;
    const long n1 = (long)d.NY * d.NZ, n2 = (long)d.NX * d.NZ, n3 = (long)d.NX * d.NY;
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (e < n1) {
        i = d.NX - 1; j = (int)(e % d.NY); k = (int)(e / d.NY);
    } else if ((e -= n1) < n2) {
        j = d.NY - 1; i = (int)(e % d.NX); k = (int)(e / d.NX);
    } else if ((e -= n2) < n3) {
        k = d.NZ - 1; i = (int)(e % d.NX); j = (int)(e / d.NX);
    } else {
        return;
    }
    const long id = gidx(d, i, j, k);
    if (H.st[id] == TB_IMMERSED) return;   // kept, as the sweep keeps IMMERSED rows
    const double v = b[id] / (H.ap_solid + H.mod);
    x1[id] = v;
    x2[id] = v;
}

// IMMERSED rows, Jacobi over the IB table: new value of row a into tmp[a]
// (x_im = (b_im - sum_{g != im} v_g x_g) / v_im), then scattered, so the
// result does not depend on the order in which rows run
__global__ void k_ib_gs(const long *__restrict__ row, const int *__restrict__ cnt, const long *__restrict__ col,
                        const double *__restrict__ val, int n, const double *__restrict__ x,
                        const double *__restrict__ b, double *__restrict__ tmp) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const long r = row[a];
    const int m = cnt[a];
    // every column and value, then every gather, issued before the first
    // is used; the sum runs in column order as a loop over them would
    long ci[9];
    double cv[9], xv[9];
#pragma unroll
    for (int e = 0; e < 9; e++) {
        ci[e] = e < m ? col[(long)a * 9 + e] : r;
        cv[e] = e < m ? val[(long)a * 9 + e] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 9; e++) xv[e] = (e < m && ci[e] != r) ? x[ci[e]] : 0.0;
    double s = 0.0, dg = 1.0;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        if (e >= m) continue;
        if (ci[e] == r) dg = cv[e];
        else s += cv[e] * xv[e];
    }
    tmp[a] = (b[r] - s) / dg;
}

__global__ void k_ib_scatter(const long *__restrict__ row, int n, const double *__restrict__ tmp,
                             double *__restrict__ x) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) x[row[a]] = tmp[a];
}

// ||b - A x||^2 and ||b||^2 into S_TMP1, S_TMP2 (host mirror)
// |b - A x|^2 and |b|^2; with x0 given also the increment x - x0 into dn
static int helm_residual(tb_ctx *c, const HelmOp &H, const double *x, const double *b, double *ibuf, double &rr,
                         double &bb, const double *x0 = nullptr, float *dn = nullptr) {
    int nb = helm_apply(c, H, x, ibuf, false, x0 ? 4 : 3, b, x0, dn);
    if (nb < 0 || tb_reduce_finish(c, nb, 2, S_TMP1)) return -1;
    if (tb_read_scalars(c, S_TMP1, 2)) return -1;
    rr = c->hscal[S_TMP1];
    bb = c->hscal[S_TMP2];
    return 0;
}

// n sweep pairs, each followed by the Jacobi update of the IMMERSED rows,
// alternating between the buffers cur and alt (the same size and pitch);
// cur is swapped to the buffer that holds the result.  With src given the
// first pair starts from src (plus inc, when given) instead of cur; inc may
// be cur itself: the last planes written into cur first are read there only
// with zero coefficients (no row couples to them).
static int gs_pairs(tb_ctx *c, const HelmOp &H, double *&cur, double *&alt, const double *b, int n,
                    const double *src = nullptr, const float *inc = nullptr) {
    const int q = H.q;
    double *tmp = c->kv[4];
    const GsTiles T = make_gs_tiles(c->d);
    if (c->d.ntot >= (1L << 31)) {
        c->err = "Helmholtz sweeps: field exceeds 2^31 elements";
        return -1;
    }
    if (n > 0) {
        const Dims &d = c->d;
        const long np = (long)d.NY * d.NZ + (long)d.NX * d.NZ + (long)d.NX * d.NY;
        hipLaunchKernelGGL(k_gs_last_planes, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->s, H, b, cur, alt);
        TB_LAUNCH_CHECK(c);
    }
    for (int sw = 0; sw < n; sw++) {
        const double *from = (sw == 0 && src) ? src : cur;
        if (sw == 0 && src && inc) {
            TB_PROF(c, TB_K_HELM_GS,
                    hipLaunchKernelGGL(k_helm_gs_pair<true>, dim3(T.per_xcd * NXCD), dim3(GS_THREADS), 0, c->s, H,
                                       from, alt, b, T, inc));
        } else {
            TB_PROF(c, TB_K_HELM_GS,
                    hipLaunchKernelGGL(k_helm_gs_pair<false>, dim3(T.per_xcd * NXCD), dim3(GS_THREADS), 0, c->s, H,
                                       from, alt, b, T, (const float *)nullptr));
        }
        double *w = cur;
        cur = alt;
        alt = w;
        if (c->nib[q] > 0) {
            const dim3 gi((c->nib[q] + 255) / 256);
            if (c->prof_on) tb_prof_begin(c, TB_K_IB);
            hipLaunchKernelGGL(k_ib_gs, gi, dim3(256), 0, c->s, c->ib_row[q], c->ib_cnt[q], c->ib_col[q],
                               c->ib_val[q], c->nib[q], (const double *)cur, b, tmp);
            hipLaunchKernelGGL(k_ib_scatter, gi, dim3(256), 0, c->s, c->ib_row[q], c->nib[q], (const double *)tmp,
                               cur);
            if (c->prof_on) tb_prof_end(c, TB_K_IB);
        }
        TB_LAUNCH_CHECK(c);
    }
    return 0;
}

// copy the box points of src into dst (row padding untouched)
static int copy_box(tb_ctx *c, double *dst, const double *src) {
    const size_t pitch = sizeof(double) * (size_t)c->d.px;
    TB_CHECK(c, hipMemcpy2DAsync(dst, pitch, src, pitch, sizeof(double) * (size_t)c->d.NX,
                                 (size_t)c->d.NY * c->d.NZ, hipMemcpyDeviceToDevice, c->s));
    return 0;
}

// returns 0 and sets ok when converged (the field of quantity q then holds
// the solution); ok = false when the sweeps stall (contraction per pair worse
// than 0.5) or run out of sweeps, x then holds the last iterate.  The initial
// guess is the field plus the increment of the quantity's last solve, when
// that solve converged on the same trajectory (the reference starts from the
// field alone; both stop at the same residual test).  Three FP64 buffers
// take part: the field x, read by the first pair and kept so that each
// residual check also writes the increment, and the two iterate buffers,
// kv[6] and the quantity's own second buffer.  On convergence the field
// takes the result's buffer, the second buffer the other iterate's and
// kv[6] the old field's (no kernel or graph keeps any of these pointers
// between calls).  The increment is kept in FP32: it only moves the start
// of the next solve, which converges to the same residual test, and the
// check that writes it and the first pair that reads it stream half the
// bytes for it.
// The true residual costs about two thirds of a sweep pair, so it is
// evaluated only where the solve may have converged: first after as many
// pairs as the last solve of the same quantity and RK stage parity needed
// (the count is steady from step to step), then after the pairs predicted
// from the observed contraction (at most 3 at a time: the contraction
// improves as the sweeps go, a longer prediction overshoots).  The initial
// residual is not evaluated (res0 = NaN).
static int helm_gs_solve(tb_ctx *c, const HelmOp &H, double *x, const double *b, double rtol, int maxit, int &its,
                         double &res0, double &res, double &bn, bool &ok) {
    const int q = H.q;
    double *ibuf = c->kv[5];
    double rr, bb;
    res0 = NAN;
    res = NAN;
    its = 0;
    ok = false;
    if (maxit < 1) return 0;
    if (tb_helm_inc_alloc(c, q)) return -1;
    HelmHistory &h = c->helm[q];
    const bool field = x == c->f[TB_F_U + q];
    const float *inc = (field && h.inc_ok) ? h.inc : nullptr;
    h.inc_ok = false;   // the increment is overwritten by the first residual check
    double *cur = h.buf, *alt = c->kv[6];
    const int par = (int)(h.nsolve++ & 1u);
    int chunk = h.sweeps[par] > 0 ? h.sweeps[par] : 2;
    double prev = NAN, rate = h.rate[par];
    const double *src = x;
    while (!ok && its < maxit) {
        const int m = chunk < maxit - its ? chunk : maxit - its;
        if (gs_pairs(c, H, cur, alt, b, m, src, inc)) return -1;
        src = nullptr;
        inc = nullptr;
        its += m;
        if (helm_residual(c, H, cur, b, ibuf, rr, bb, field ? x : nullptr, field ? h.inc : nullptr)) return -1;
        bn = sqrt(bb);
        if (bn == 0.0) bn = 1.0;
        res = sqrt(rr) / bn;
        if (!(res == res)) break;
        ok = res <= rtol;
        if (ok) break;
        if (prev == prev) rate = pow(res / prev, 1.0 / m);
        prev = res;
        if (!(rate > 0.0)) rate = 0.05;   // no estimate yet
        if (!(rate < 0.5)) break;
        const int need = (int)ceil(log(rtol / res) / log(rate));
        chunk = need < 1 ? 1 : (need > 3 ? 3 : need);
    }
    if (ok) {
        // one pair fewer next time when this residual shows it would have
        // sufficed at the last contraction
        const double r1 = (rate > 0.0 && rate < 0.5) ? rate : 0.05;
        h.sweeps[par] = (its > 1 && res / r1 <= rtol) ? its - 1 : its;
        if (prev == prev) h.rate[par] = rate;
    }
    if (ok && field) {
        // cur: the solution, h.inc: its increment over x (the iterates never
        // write x, which is now free)
        c->f[TB_F_U + q] = cur;
        h.buf = alt;
        c->kv[6] = x;
        h.inc_ok = true;
    } else if (copy_box(c, x, cur)) {
        return -1;   // BiCGStab continues from x
    }
    return 0;
}

// The smoother of tb_helmholtz_solve by itself: f_x <- `pairs` red-black
// sweep pairs (each followed by the Jacobi update of the IMMERSED rows) of
// A_q(dt) x = f_b, in place.
extern "C" int tb_helmholtz_gs_sweeps(tb_ctx *c, int q, double dt, int f_b, int f_x, int pairs) {
    if (!c || q < 0 || q > 3 || f_b < 0 || f_b >= TB_F_COUNT || f_x < 0 || f_x >= TB_F_COUNT || f_b == f_x ||
        pairs < 0) {
        if (c) c->err = "tb_helmholtz_gs_sweeps: bad argument";
        return -1;
    }
    tb_field_written(c, f_x);
    HelmOp H = make_op(c, q, dt);
    double *cur = c->f[f_x], *alt = c->kv[6];
    if (gs_pairs(c, H, cur, alt, c->f[f_b], pairs)) return -1;
    if (cur != c->f[f_x] && copy_box(c, c->f[f_x], cur)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_helmholtz_apply(tb_ctx *c, int q, double dt, int f_in, int f_out) {
    if (!c || q < 0 || q > 3 || f_in < 0 || f_in >= TB_F_COUNT || f_out < 0 || f_out >= TB_F_COUNT || f_in == f_out) {
        if (c) c->err = "tb_helmholtz_apply: bad argument";
        return -1;
    }
    tb_field_written(c, f_out);
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

static int read_scalars(tb_ctx *c, int first, int count) { return tb_read_scalars(c, first, count); }

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
    // 1 when the Gauss-Seidel sweeps stopped contracting (or ran out of
    // pairs) and BiCGStab finished the solve
    int handover = 0;
    if (c->helm_method == TB_HELM_GS) {
        int gits = 0;
        double g0 = 0.0, gres = 0.0, gbn = 1.0;
        bool ok = false;
        if (helm_gs_solve(c, H, x, b, rtol, maxit, gits, g0, gres, gbn, ok)) return -1;
        if (ok) {
            TB_CHECK(c, hipEventRecord(c->ev1, c->s));
            if (tb_sync(c)) return -1;
            if (st) {
                st->iterations = gits;
                st->converged = 1;
                st->fallback = 0;
                st->vcycles = 0;
                st->res0 = g0;
                st->res = gres;
                st->ms = tb_elapsed_ms(c);
            }
            return 0;
        }
        // sweeps stalled: BiCGStab continues from the current x
        handover = 1;
    }
    c->helm[q].inc_ok = false;   // the increment is recorded by the Gauss-Seidel solve only
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
        st->fallback = handover;
        st->vcycles = 0;
        st->res0 = res0;
        st->res = tres;
        st->ms = tb_elapsed_ms(c);
    }
    return (tres <= rtol) ? 0 : 1;
}
