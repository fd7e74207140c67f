// tb_eno.hip -- ENO convective terms for u, v, w and c, and the
// right-hand sides of the four Helmholtz systems.
//
// The reference evaluates ENO_compute_convective_terms (ENO.c:103-245) one
// grid line at a time after gathering the line and its wall ghosts into 1-D
// buffers.  Here every thread evaluates one node directly from the field in
// HBM: it reads the +-order window along each axis, substitutes the wall
// ghost values of Velocity_*_ENO_ghost_cells / Conc_ENO_ghost_cells where the
// window leaves the box, and applies the same divided differences and
// stencil choice.  The advecting velocities of the transposes
// (Velocity_transpose_velocities, Velocity.c:3341-3663) and of the cell
// centres (Velocity_cell_center, Velocity.c:1476-1602) are interpolated in
// registers.  All arithmetic keeps the reference's order (no contraction),
// so results are bit-identical to the line-by-line evaluation.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "tb_common.h"

// MyMath_interpolate_quantity, four points (MyMath.c:52-73)
__device__ inline double interp4d(double q0, double q1, double q2, double q3, double X10, double X11, double X20,
                                  double X22, double x1, double x2) {
    double a_Dx = 1.0 / (X11 - X10);
    double Wp0 = fabs((x1 - X10) * a_Dx);
    double Wp1 = 1.0 - Wp0;
    double q_top = Wp0 * q2 + Wp1 * q3;
    double q_bottom = Wp0 * q0 + Wp1 * q1;
    double a_Dy = 1.0 / (X22 - X20);
    Wp0 = fabs((x2 - X20) * a_Dy);
    double Wp2 = 1.0 - Wp0;
    return Wp0 * q_bottom + Wp2 * q_top;
}

// interp4d with its two weights given: Wa = |(x1 - X10) / (X11 - X10)| and
// Wb = |(x2 - X20) / (X22 - X20)| depend on one grid index each, so they come
// from tables built once with the same operations (1/(X11 - X10) first, then
// the product), which keeps the result bit-identical
__device__ inline double interp4w(double q0, double q1, double q2, double q3, double Wa, double Wb) {
    double Wp1 = 1.0 - Wa;
    double q_top = Wa * q2 + Wp1 * q3;
    double q_bottom = Wa * q0 + Wp1 * q1;
    double Wp2 = 1.0 - Wb;
    return Wb * q_bottom + Wp2 * q_top;
}

// the twelve weight tables of the transposes, one per (interpolation, axis)
enum { W_VU_X, W_VU_Y, W_WU_X, W_WU_Z, W_UV_X, W_UV_Y, W_WV_Z, W_WV_Y, W_UW_X, W_UW_Z, W_VW_Z, W_VW_Y, W_COUNT };

// One line direction of one quantity: interior span and wall ghost rules.
struct LineSpec {
    int lo, hi;            // interior (non-BOUNDARY) range along the axis
    int kind;              // 0: normal component on the wall; 1: mirrored line
    int ux_quirk;          // u along x: left mirrors take -D0[order] (Velocity.c:2191)
    int hi_mode;           // 1: top value scaled by hi_scale (particle no-flux, Conc.c:1164-1179)
    double wall_lo, wall_hi;
    double sgn_lo, sgn_hi; // kind 1: -1 negate (no-slip), +1 copy (free-slip / Neumann)
    double hi_scale;
    const double *pos;     // coordinates of the quantity's nodes along the axis
    long stride;           // element stride of the axis in the padded layout
    int uni;               // 1: every gap pos[m+1]-pos[m] of [lo, hi] is h, every pos[m+1]-pos[m-1] is 2h,
                           //    h a power of two (pow2_uniform)
    double h, rh;          // that h and its exact reciprocal
    // x lines next to an open face: every ghost of that side takes the
    // plane's value at the line's (j, k) (null: the wall rule)
    const double *open_lo, *open_hi;
};

// Window of 2R+1 values and positions around axis position m0 (buffer
// index R + m0 - lo of the reference's line).
template <int R>
__device__ inline void load_window(const LineSpec &L, const double *__restrict__ f, long base, int m0,
                                   double (&w)[7], double (&p)[7], double olo, double ohi) {
    // every load first, at axis positions clamped into the interior span, so
    // the loads are issued together; a position left of lo therefore holds
    // the first interior value D0[order], one right of hi the last one
#pragma unroll
    for (int o = -R; o <= R; o++) {
        int m = m0 + o;
        int mc = m < L.lo ? L.lo : (m > L.hi ? L.hi : m);
        w[3 + o] = f[base + (long)(mc - m0) * L.stride];
        p[3 + o] = L.pos[mc];
    }
    if (m0 - R >= L.lo && m0 + R <= L.hi) return;
    // wall ghosts where the window leaves the box
#pragma unroll
    for (int o = -R; o <= R; o++) {
        int m = m0 + o;
        if (m < L.lo) {
            int ex = R + m - L.lo;                  // ghost_index
            double first = w[3 + o];
            double wv, pv;
            if (L.kind == 0) {
                if (ex == R - 1) { pv = L.wall_lo; wv = 0.0; }
                else { pv = 2.0 * L.wall_lo - L.pos[L.lo + R - ex - 2]; wv = L.ux_quirk ? -first : 0.0; }
            } else {
                pv = 2.0 * L.wall_lo - L.pos[L.lo + R - ex - 1];
                wv = (L.sgn_lo < 0.0) ? -first : +first;
            }
            if (L.open_lo) wv = olo;
            w[3 + o] = wv;
            p[3 + o] = pv;
        } else if (m > L.hi) {
            int n = L.hi - L.lo + 1;
            int b = R + m - L.lo;
            int ex = n + 2 * R - 1 - b;
            double last = w[3 + o];
            double wv, pv;
            if (L.kind == 0) {
                if (ex == R - 1) { pv = L.wall_hi; wv = 0.0; }
                else { pv = 2.0 * L.wall_hi - L.pos[L.hi + ex + 2 - R]; wv = 0.0; }
            } else {
                pv = 2.0 * L.wall_hi - L.pos[L.hi + 1 + ex - R];
                if (L.hi_mode) wv = L.hi_scale * last;
                else wv = (L.sgn_hi < 0.0) ? -last : +last;
            }
            if (L.open_hi) wv = ohi;
            w[3 + o] = wv;
            p[3 + o] = pv;
        }
    }
}

// ENO_compute_convective_terms for the node at window centre 3.
template <int R>
__device__ inline double eno_eval(const double (&w)[7], const double (&p)[7], double vel) {
    // first divided differences D1[m] = (D0[m+1]-D0[m])/(x[m+1]-x[m]), m = -R..R-1
    double D1[6];
#pragma unroll
    for (int m = -R; m < R; m++) D1[m + 3] = (w[3 + m + 1] - w[3 + m]) / (p[3 + m + 1] - p[3 + m]);
    bool left = vel > 0.0;                     // k = i-1 if V > 0 else i
    double dQ = left ? D1[2] : D1[3];
    if (R > 1) {
        // D2[m] = (D1[m]-D1[m-1])/(x[m+1]-x[m-1]) for m = -R+1..R-1
        double D2[5];
#pragma unroll
        for (int m = -R + 1; m <= R - 1; m++)
            D2[m + 2] = (D1[m + 3] - D1[m + 2]) / (p[3 + m + 1] - p[3 + m - 1]);
        // k = -1 or 0
        double d2k = left ? D2[1] : D2[2];
        double d2k1 = left ? D2[2] : D2[3];
        double c;
        int kstar;
        int k = left ? -1 : 0;
        if (fabs(d2k) <= fabs(d2k1)) { c = d2k; kstar = k - 1; }
        else { c = d2k1; kstar = k; }
        double x_i = p[3];
        double x_k = left ? p[2] : p[3];
        double x_k1 = left ? p[3] : p[4];
        dQ += c * ((x_i - x_k) + (x_i - x_k1));
        if (R > 2) {
            // D3[m] = (D2[m+1]-D2[m])/(x[m+2]-x[m-1]) for m = -2..1
            double D3[4];
#pragma unroll
            for (int m = -2; m <= 1; m++)
                D3[m + 2] = (D2[m + 3] - D2[m + 2]) / (p[3 + m + 2] - p[3 + m - 1]);
            // kstar in {-2,-1,0}
            double d3a = (kstar == -2) ? D3[0] : (kstar == -1) ? D3[1] : D3[2];
            double d3b = (kstar == -2) ? D3[1] : (kstar == -1) ? D3[2] : D3[3];
            double cstar = (fabs(d3a) <= fabs(d3b)) ? d3a : d3b;
            double x_ks = (kstar == -2) ? p[1] : (kstar == -1) ? p[2] : p[3];
            double x_ks1 = (kstar == -2) ? p[2] : (kstar == -1) ? p[3] : p[4];
            double x_ks2 = (kstar == -2) ? p[3] : (kstar == -1) ? p[4] : p[5];
            dQ += cstar * ((x_i - x_ks) * (x_i - x_ks1) + (x_i - x_ks1) * (x_i - x_ks2) + (x_i - x_ks) * (x_i - x_ks2));
        }
    }
    return vel * dQ;
}

// eno_eval<R> (R <= 2) for a window inside [lo, hi] of a uniform axis with a
// power-of-two spacing h: a quotient by h or 2h is the product with its exact
// reciprocal (both are the correctly rounded a/h), and the stencil term's
// position factor (x_i - x_k) + (x_i - x_k1) evaluates to +h (V > 0) or -h,
// so the result is bit-identical to eno_eval<R> on the loaded positions.
template <int R>
__device__ inline double eno_uniform(const LineSpec &L, const double *__restrict__ f, long base, double vel) {
    double w[5];
#pragma unroll
    for (int o = -R; o <= R; o++) w[2 + o] = f[base + (long)o * L.stride];
    const double rh = L.rh;
    double D1[4];   // D1[m + 2] = (w[m+1] - w[m]) / h, m = -R..R-1
#pragma unroll
    for (int m = -R; m < R; m++) D1[m + 2] = (w[2 + m + 1] - w[2 + m]) * rh;
    bool left = vel > 0.0;
    double dQ = left ? D1[1] : D1[2];
    if (R > 1) {
        const double r2h = 0.5 * rh;
        double D2[3];   // D2[m + 1] = (D1[m] - D1[m-1]) / 2h, m = -1..1
#pragma unroll
        for (int m = -1; m <= 1; m++) D2[m + 1] = (D1[m + 2] - D1[m + 1]) * r2h;
        double d2k = left ? D2[0] : D2[1];
        double d2k1 = left ? D2[1] : D2[2];
        double c = (fabs(d2k) <= fabs(d2k1)) ? d2k : d2k1;
        dQ += c * (left ? L.h : -L.h);
    }
    return vel * dQ;
}

// pidx: the node's (j, k) plane entry k NY + j (x lines; read only where the
// window reaches an open face)
template <int R>
__device__ inline double eno_line_point(const LineSpec &L, const double *__restrict__ f, long base, int m0,
                                        double vel, long pidx = 0) {
    if constexpr (R <= 2) {
        if (L.uni && m0 - R >= L.lo && m0 + R <= L.hi) return eno_uniform<R>(L, f, base, vel);
    }
    double w[7], p[7];
    const double olo = (L.open_lo && m0 - R < L.lo) ? L.open_lo[pidx] : 0.0;
    const double ohi = (L.open_hi && m0 + R > L.hi) ? L.open_hi[pidx] : 0.0;
    load_window<R>(L, f, base, m0, w, p, olo, ohi);
    return eno_eval<R>(w, p, vel);
}

// Per-quantity line specs for the three axes.
struct QLines {
    LineSpec ax[3];
};

struct MomArgs {
    Dims d;
    GridDev g;
    const double *W;       // W_COUNT tables of Wmax entries
    int Wmax;
    const double *u, *v, *w;
    const uint8_t *su, *sv, *sw;
    double *cu, *cv, *cw;
    QLines L[3];
    // right-hand sides in the same pass (RHS = true): the k_rhs_vel formulas
    const double *dpx, *dpy, *dpz, *cfld;
    double *ru, *rv, *rw;
    double a_dt, buoy;
    int store_conv;   // 0 with RHS: the convective terms are consumed in registers only
    const double *uin, *uout;   // u's planes of open x faces (null: closed)
};

__device__ inline bool vel_row(int s) { return s == TB_FLUID || s == TB_BUFFER_SOLID; }

template <int R, bool RHS>
__device__ inline void mom_point(const MomArgs &a, int i, int j, int k) {
    const Dims &d = a.d;
    const GridDev &g = a.g;
    long id = gidx(d, i, j, k);
    const double *u = a.u, *v = a.v, *w = a.w;
    const double *W = a.W;
    const int wm = a.Wmax;
    const long pidx = (long)k * d.NY + j;
    // u node
    {
        double r = 0.0;
        const int s = a.su[id];
        if (s != TB_BOUNDARY) {
            double vu = interp4w(v[id - 1], v[id], v[id + d.px - 1], v[id + d.px], W[W_VU_X * wm + i],
                                 W[W_VU_Y * wm + j]);
            double wu = interp4w(w[id - 1], w[id], w[id + d.plane - 1], w[id + d.plane], W[W_WU_X * wm + i],
                                 W[W_WU_Z * wm + k]);
            r = eno_line_point<R>(a.L[0].ax[0], u, id, i, u[id], pidx);
            r += eno_line_point<R>(a.L[0].ax[1], u, id, j, vu);
            r += eno_line_point<R>(a.L[0].ax[2], u, id, k, wu);
        }
        if (!RHS || a.store_conv) st_stream(&a.cu[id], r);
        if constexpr (RHS) st_stream(&a.ru[id], vel_row(s) ? u[id] * a.a_dt - r - a.dpx[id] : 0.0);
    }
    // v node
    {
        double r = 0.0;
        const int s = a.sv[id];
        if (s != TB_BOUNDARY) {
            // u beside an open face from its plane (Velocity.c:3491-3506)
            double u0 = u[id - d.px], u1 = u[id - d.px + 1], u2 = u[id], u3 = u[id + 1];
            if (a.uin && i == 0) { u0 = a.uin[pidx - 1]; u2 = a.uin[pidx]; }
            if (a.uout && i == d.NX - 2) { u1 = a.uout[pidx - 1]; u3 = a.uout[pidx]; }
            double uv = interp4w(u0, u1, u2, u3, W[W_UV_X * wm + i], W[W_UV_Y * wm + j]);
            double wv = interp4w(w[id - d.px], w[id + d.plane - d.px], w[id], w[id + d.plane], W[W_WV_Z * wm + k],
                                 W[W_WV_Y * wm + j]);
            r = eno_line_point<R>(a.L[1].ax[0], v, id, i, uv, pidx);
            r += eno_line_point<R>(a.L[1].ax[1], v, id, j, v[id]);
            r += eno_line_point<R>(a.L[1].ax[2], v, id, k, wv);
        }
        if (!RHS || a.store_conv) st_stream(&a.cv[id], r);
        if constexpr (RHS) {
            double rh = 0.0;
            if (vel_row(s)) {
                const double c_term = interp2d(a.cfld[id - d.px], g.yc[j - 1], a.cfld[id], g.yc[j], g.yv[j]);
                rh = v[id] * a.a_dt - r - a.dpy[id] - a.buoy * c_term;
            }
            st_stream(&a.rv[id], rh);
        }
    }
    // w node
    {
        double r = 0.0;
        const int s = a.sw[id];
        if (s != TB_BOUNDARY) {
            // u beside an open face from its plane (Velocity.c:3589-3603)
            double u0 = u[id - d.plane], u1 = u[id - d.plane + 1], u2 = u[id], u3 = u[id + 1];
            if (a.uin && i == 0) { u0 = a.uin[pidx - d.NY]; u2 = a.uin[pidx]; }
            if (a.uout && i == d.NX - 2) { u1 = a.uout[pidx - d.NY]; u3 = a.uout[pidx]; }
            double uw = interp4w(u0, u1, u2, u3, W[W_UW_X * wm + i], W[W_UW_Z * wm + k]);
            double vw = interp4w(v[id - d.plane], v[id], v[id - d.plane + d.px], v[id + d.px], W[W_VW_Z * wm + k],
                                 W[W_VW_Y * wm + j]);
            r = eno_line_point<R>(a.L[2].ax[0], w, id, i, uw, pidx);
            r += eno_line_point<R>(a.L[2].ax[1], w, id, j, vw);
            r += eno_line_point<R>(a.L[2].ax[2], w, id, k, w[id]);
        }
        if (!RHS || a.store_conv) st_stream(&a.cw[id], r);
        if constexpr (RHS) st_stream(&a.rw[id], vel_row(s) ? w[id] * a.a_dt - r - a.dpz[id] : 0.0);
    }
}

template <int R, bool RHS>
__global__ void __launch_bounds__(256) k_eno_momentum(MomArgs a, Tiles t) {
    int i, j, k;
    if (!tile_point_kfast(t, i, j, k)) return;
    mom_point<R, RHS>(a, i, j, k);
}

struct ConcArgs {
    Dims d;
    GridDev g;
    const double *u, *v, *w, *c;
    const uint8_t *sc;
    double *cc;
    double v_settle;
    int particle;
    int hindered;   // particle and params->hindered_settling
    double phi_max;
    QLines L;
    double *rc;     // right-hand side in the same pass (RHS = true): the k_rhs_conc formula
    double a_dt;
    int store_conv;
    const double *uin, *uout;   // u's planes of open x faces (null: closed)
};

template <int R, bool RHS>
__device__ inline void conc_point(const ConcArgs &a, int i, int j, int k) {
    const Dims &d = a.d;
    const GridDev &g = a.g;
    long id = gidx(d, i, j, k);
    double r = 0.0;
    const int s = a.sc[id];
    if (s != TB_BOUNDARY) {
        // Velocity_cell_center (Velocity.c:1476-1602) at this cell; u on an
        // open face from its plane (Velocity.c:1544-1558)
        const long pidx = (long)k * d.NY + j;
        const double uw_ = (a.uin && i == 0) ? a.uin[pidx] : a.u[id];
        const double ue_ = (a.uout && i == d.NX - 2) ? a.uout[pidx] : a.u[id + 1];
        double ubc = interp2d(uw_, g.xu[i], ue_, g.xu[i + 1], g.xc[i]);
        double vbc = interp2d(a.v[id], g.yv[j], a.v[id + d.px], g.yv[j + 1], g.yc[j]);
        double wbc = interp2d(a.w[id], g.zw[k], a.w[id + d.plane], g.zw[k + 1], g.zc[k]);
        if (a.particle) vbc = particle_v(vbc, s, a.c[id], a.hindered, a.phi_max, a.v_settle);
        r = eno_line_point<R>(a.L.ax[0], a.c, id, i, ubc, pidx);
        r += eno_line_point<R>(a.L.ax[1], a.c, id, j, vbc);
        r += eno_line_point<R>(a.L.ax[2], a.c, id, k, wbc);
    }
    if (!RHS || a.store_conv) st_stream(&a.cc[id], r);
    if constexpr (RHS) st_stream(&a.rc[id], s == TB_FLUID ? a.c[id] * a.a_dt - r : 0.0);
}

template <int R, bool RHS>
__global__ void __launch_bounds__(256) k_eno_conc(ConcArgs a, Tiles t) {
    int i, j, k;
    if (!tile_point_kfast(t, i, j, k)) return;
    conc_point<R, RHS>(a, i, j, k);
}

// True when every first gap fl(p[m+1] - p[m]) of [lo, hi] equals one power
// of two h and every second gap fl(p[m+1] - p[m-1]) equals 2h, the exact
// divisors of eno_eval<R <= 2> in a window inside [lo, hi].
static bool pow2_uniform(const std::vector<double> &p, int lo, int hi, double &h) {
    if (lo < 0 || hi >= (int)p.size() || hi - lo < 2) return false;
    h = p[lo + 1] - p[lo];
    int e;
    if (!(h > 0.0) || !std::isfinite(h) || std::frexp(h, &e) != 0.5 || !std::isnormal(1.0 / h)) return false;
    for (int m = lo; m < hi; m++)
        if (p[m + 1] - p[m] != h) return false;
    for (int m = lo + 1; m < hi; m++)
        if (p[m + 1] - p[m - 1] != 2.0 * h) return false;
    return true;
}

// Line specs per quantity (the ghost rules of Velocity.c:2124-2801 and
// Conc.c:1021-1220; wall order x0,x1,y0,y1,z0,z1).
static QLines make_lines(tb_ctx *c, int q) {
    QLines Q;
    const Dims &d = c->d;
    const double *posq[4][3] = {{c->g.xu, c->g.yu, c->g.zu},
                                {c->g.xv, c->g.yv, c->g.zv},
                                {c->g.xw, c->g.yw, c->g.zw},
                                {c->g.xc, c->g.yc, c->g.zc}};
    int hi[3] = {d.NX - 2, d.NY - 2, d.NZ - 2};
    double wlo[3] = {c->hgrid[3][0], c->hgrid[7][0], c->hgrid[11][0]};                  // xu[0], yv[0], zw[0]
    double whi[3] = {c->hgrid[3][d.NX - 1], c->hgrid[7][d.NY - 1], c->hgrid[11][d.NZ - 1]};
    long stride[3] = {1, d.px, d.plane};
    for (int ax = 0; ax < 3; ax++) {
        LineSpec &L = Q.ax[ax];
        L.lo = (q == ax) ? 1 : 0;
        L.hi = hi[ax];
        L.kind = (q == ax) ? 0 : 1;
        L.ux_quirk = (q == TB_Q_U && ax == 0);
        L.hi_mode = 0;
        L.hi_scale = 1.0;
        L.wall_lo = wlo[ax];
        L.wall_hi = whi[ax];
        L.pos = posq[q][ax];
        L.stride = stride[ax];
        // open x faces (Velocity.c:2193-2233, 2442-2444, 2669-2671; Conc.c:1099-1122)
        L.open_lo = (ax == 0 && (q == TB_Q_U || q == TB_Q_C)) ? tb_in_plane(c, q) : nullptr;
        L.open_hi = ax == 0 ? tb_out_plane(c, q) : nullptr;
        const int hq = (q == TB_Q_C) ? ax : 3 + 3 * q + ax;   // host copy of posq[q][ax]
        L.uni = pow2_uniform(c->hgrid[hq], L.lo, L.hi, L.h) ? 1 : 0;
        L.rh = L.uni ? 1.0 / L.h : 0.0;
        if (!L.uni) L.h = 0.0;
        if (q != TB_Q_C) {
            L.sgn_lo = c->cfg.wall_noslip[2 * ax] ? -1.0 : 1.0;
            L.sgn_hi = c->cfg.wall_noslip[2 * ax + 1] ? -1.0 : 1.0;
        } else {
            L.sgn_lo = 1.0;
            L.sgn_hi = 1.0;
            if (ax == 1 && c->cfg.conc_type == TB_PARTICLE && c->cfg.sedimentation) {
                double coef = 2.0 / (c->cfg.pe * c->dXi) * c->hgrid[19][d.NY - 1];
                double Vs = fabs(c->cfg.v_settle);
                L.hi_mode = 1;
                L.hi_scale = (coef - Vs) / (coef + Vs);
            }
        }
    }
    return Q;
}

// weight Wa = |(x1 - X10) * (1 / (X11 - X10))| of interp4d, 0 where an
// index leaves the grid (those nodes are BOUNDARY and never interpolated)
static double iw(const std::vector<double> &x1, int i1, const std::vector<double> &X, int i0, int i11) {
    const int n = (int)X.size();
    if (i1 < 0 || i1 >= (int)x1.size() || i0 < 0 || i0 >= n || i11 < 0 || i11 >= n) return 0.0;
    const double aD = 1.0 / (X[i11] - X[i0]);
    return fabs((x1[i1] - X[i0]) * aD);
}

static int eno_weights(tb_ctx *c) {
    if (c->eno_w) return 0;
    const auto &xu = c->hgrid[3], &yu = c->hgrid[4], &zu = c->hgrid[5];
    const auto &xv = c->hgrid[6], &yv = c->hgrid[7], &zv = c->hgrid[8];
    const auto &xw = c->hgrid[9], &yw = c->hgrid[10], &zw = c->hgrid[11];
    const int wm = std::max(c->d.NX, std::max(c->d.NY, c->d.NZ));
    std::vector<double> W((size_t)W_COUNT * wm, 0.0);
    for (int i = 0; i < c->d.NX; i++) {
        W[W_VU_X * wm + i] = iw(xu, i, xv, i - 1, i);
        W[W_WU_X * wm + i] = iw(xu, i, xw, i - 1, i);
        W[W_UV_X * wm + i] = iw(xv, i, xu, i, i + 1);
        W[W_UW_X * wm + i] = iw(xw, i, xu, i, i + 1);
    }
    for (int j = 0; j < c->d.NY; j++) {
        W[W_VU_Y * wm + j] = iw(yu, j, yv, j, j + 1);
        W[W_UV_Y * wm + j] = iw(yv, j, yu, j - 1, j);
        W[W_WV_Y * wm + j] = iw(yv, j, yw, j - 1, j);
        W[W_VW_Y * wm + j] = iw(yw, j, yv, j, j + 1);
    }
    for (int k = 0; k < c->d.NZ; k++) {
        W[W_WU_Z * wm + k] = iw(zu, k, zw, k, k + 1);
        W[W_WV_Z * wm + k] = iw(zv, k, zw, k, k + 1);
        W[W_UW_Z * wm + k] = iw(zw, k, zu, k - 1, k);
        W[W_VW_Z * wm + k] = iw(zw, k, zv, k - 1, k);
    }
    TB_CHECK(c, hipMalloc((void **)&c->eno_w, sizeof(double) * W.size()));
    TB_CHECK(c, hipMemcpy(c->eno_w, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice));
    c->eno_wmax = wm;
    return 0;
}

// ---------------------------------------------------------------------------
// Second-order ENO of u, v, w and c in one k-marching pass (the workloads'
// order) on power-of-two uniform grids.  A block owns a 64 x 4 tile of (i, j)
// columns and walks EF_KC planes of it.  The planes k-1, k, k+1 of u, v, w, c
// sit in an LDS ring with a two-point x / y halo, so the x and y windows, the
// z neighbours k +- 1 and the transposes read LDS; each thread keeps k - 2 and
// k + 2 of its own column in registers.  It computes the nodes at least
// EF_SHELL from every box face, whose windows all lie inside the uniform
// interior (the exact-reciprocal form of eno_uniform); k_eno_shell computes
// the rest with the per-point code (wall ghosts, general form).  Both do the
// operations of eno_line_point in the same order, so every value is
// bit-identical to the per-point kernels.
constexpr int EF_TX = 64, EF_TY = 4, EF_H = 2, EF_KC = 16;
constexpr int EF_SX = EF_TX + 2 * EF_H, EF_SY = EF_TY + 2 * EF_H;
constexpr int EF_PLANE = EF_SX * EF_SY;
constexpr int EF_THREADS = EF_TX * EF_TY;
constexpr int EF_STAGE = (EF_PLANE + EF_THREADS - 1) / EF_THREADS;

struct EfTiles {
    int tx, ty, kc;     // tiles per axis from i, j = 0, k chunks from k = EF_SHELL
    long n, per_xcd;
};

static EfTiles make_ef_tiles(const Dims &d, int shell) {
    EfTiles t;
    t.tx = (d.NX - shell + EF_TX - 1) / EF_TX;
    t.ty = (d.NY - shell + EF_TY - 1) / EF_TY;
    t.kc = (d.NZ - 2 * shell + EF_KC - 1) / EF_KC;
    t.n = (long)t.tx * t.ty * t.kc;
    t.per_xcd = (t.n + NXCD - 1) / NXCD;
    return t;
}

// eno_uniform<2> on five values already at hand (w[2] is the node)
__device__ inline double eno_u5(double w0, double w1, double w2, double w3, double w4, double rh, double h,
                                double vel) {
    const double D1_0 = (w1 - w0) * rh, D1_1 = (w2 - w1) * rh, D1_2 = (w3 - w2) * rh, D1_3 = (w4 - w3) * rh;
    const bool left = vel > 0.0;
    double dQ = left ? D1_1 : D1_2;
    const double r2h = 0.5 * rh;
    const double D2_0 = (D1_1 - D1_0) * r2h, D2_1 = (D1_2 - D1_1) * r2h, D2_2 = (D1_3 - D1_2) * r2h;
    const double d2k = left ? D2_0 : D2_1;
    const double d2k1 = left ? D2_1 : D2_2;
    const double c = (fabs(d2k) <= fabs(d2k1)) ? d2k : d2k1;
    dQ += c * (left ? h : -h);
    return vel * dQ;
}

// Everything k_eno_fused2 reads, flat.
struct FusedArgs {
    Dims d;
    const double *u, *v, *w, *c;
    const uint8_t *su, *sv, *sw, *sc;
    const double *dpx, *dpy, *dpz;
    double *cu, *cv, *cw, *cc, *ru, *rv, *rw, *rc;
    const double *W;          // transpose weight tables (W_COUNT x wm)
    int wm;
    const double *xu, *xc, *yv, *yc, *zw, *zc;
    double a_dt, buoy, v_settle, phi_max;
    int particle, hindered, store_conv;
    // uniform fast path: bit 3*q + ax set when axis ax of quantity q is a
    // power-of-two uniform line of spacing h[ax]; hi[ax] the last interior
    // index (the first is 1 on the quantity's own axis, else 0)
    int unimask;
    int hi[3];
    double h[3], rh[3];
};

// the fused kernel's nodes: at least EF_SHELL from every box face, where
// every window of every quantity lies inside its line's interior span
constexpr int EF_SHELL = 3;

// a value every lane of the wave holds (j is one per wave: a 64-wide row of
// the 64 x 4 block), moved to scalar registers, bits unchanged
__device__ inline double wave_uniform(double v) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)bits);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(bits >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <bool RHS>
__global__ void __launch_bounds__(EF_THREADS, 3) k_eno_fused2(FusedArgs A, EfTiles T) {
    // [ring slot][field u, v, w, c][plane with halo]
    __shared__ double S[3][4][EF_PLANE];
    // per-plane values, the same for every thread, staged once: the z
    // weights of the four transposes, zw and zc (vector loads of them inside
    // the plane loop would each wait on every load issued before them)
    __shared__ double ZK[EF_KC + 1][6];
    const Dims &d = A.d;
    const long b = blockIdx.x;
    const long lt = (b % NXCD) * T.per_xcd + b / NXCD;
    if (lt >= T.n) return;                     // whole block: no barrier below is shared
    const int tyi = (int)(lt % T.ty);
    const long r1 = lt / T.ty;
    const int txi = (int)(r1 % T.tx);
    const int kci = (int)(r1 / T.tx);
    const int i0 = txi * EF_TX, j0 = tyi * EF_TY;
    const int k0 = EF_SHELL + kci * EF_KC;
    const int k1 = (k0 + EF_KC < d.NZ - EF_SHELL) ? k0 + EF_KC : d.NZ - EF_SHELL;
    for (int e = threadIdx.y * EF_TX + threadIdx.x; e < 6 * (k1 - k0 + 1); e += EF_THREADS) {
        const int m = e / 6, f = e - 6 * m, kk = k0 + m;   // kk <= k1 < NZ
        const int tab[4] = {W_WU_Z, W_WV_Z, W_UW_Z, W_VW_Z};
        ZK[m][f] = f < 4 ? A.W[tab[f] * A.wm + kk] : (f == 4 ? A.zw[kk] : A.zc[kk]);
    }
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * EF_TX + tx;
    const int i = i0 + tx, j = j0 + ty;
    const bool act = i >= EF_SHELL && i < d.NX - EF_SHELL && j >= EF_SHELL && j < d.NY - EF_SHELL;
    // weights and coordinates that do not change with k (clamped so that
    // threads outside the box read inside it; they are never used there)
    const int ix = i < d.NX - 2 ? i : d.NX - 2;
    const double wvu_x = A.W[W_VU_X * A.wm + ix], wwu_x = A.W[W_WU_X * A.wm + ix];
    const double wuv_x = A.W[W_UV_X * A.wm + ix], wuw_x = A.W[W_UW_X * A.wm + ix];
    const double xu0 = A.xu[ix], xu1 = A.xu[ix + 1], xc0 = A.xc[ix];
    const int jy = j < 1 ? 1 : (j > d.NY - 2 ? d.NY - 2 : j);
    const double wvu_y = wave_uniform(A.W[W_VU_Y * A.wm + jy]), wuv_y = wave_uniform(A.W[W_UV_Y * A.wm + jy]);
    const double wwv_y = wave_uniform(A.W[W_WV_Y * A.wm + jy]), wvw_y = wave_uniform(A.W[W_VW_Y * A.wm + jy]);
    const double yv0 = wave_uniform(A.yv[jy]), yv1 = wave_uniform(A.yv[jy + 1]);
    const double yc0 = wave_uniform(A.yc[jy]), ycm = wave_uniform(A.yc[jy - 1]);

    // staging offsets of this thread's halo points (clamped into the box;
    // a clamped point is never read by an active node)
    int soff[EF_STAGE];
#pragma unroll
    for (int e = 0; e < EF_STAGE; e++) {
        const int p = t + e * EF_THREADS;
        const int r = p / EF_SX, cc = p - r * EF_SX;
        int ii = i0 - EF_H + cc, jj = j0 - EF_H + r;
        ii = ii < 0 ? 0 : (ii > d.NX - 1 ? d.NX - 1 : ii);
        jj = jj < 0 ? 0 : (jj > d.NY - 1 ? d.NY - 1 : jj);
        soff[e] = p < EF_PLANE ? jj * d.px + ii : -1;
    }
    auto stage = [&](int m) {
        const long base = (long)m * d.plane;
        double *P = &S[(m + 3) % 3][0][0];
#pragma unroll
        for (int e = 0; e < EF_STAGE; e++) {
            const int p = t + e * EF_THREADS;
            if (soff[e] >= 0) {
                const long q = base + soff[e];
                P[p] = A.u[q];
                P[EF_PLANE + p] = A.v[q];
                P[2 * EF_PLANE + p] = A.w[q];
                P[3 * EF_PLANE + p] = A.c[q];
            }
        }
    };
    // the same in two halves: plane m into registers one iteration ahead,
    // then into its ring slot, so the loads overlap the passes between
    double pf[EF_STAGE][4];
    auto fetch = [&](int m) {
        const long base = (long)m * d.plane;
#pragma unroll
        for (int e = 0; e < EF_STAGE; e++) {
            if (soff[e] >= 0) {
                const long q = base + soff[e];
                pf[e][0] = A.u[q];
                pf[e][1] = A.v[q];
                pf[e][2] = A.w[q];
                pf[e][3] = A.c[q];
            }
        }
    };
    auto put = [&](int m) {
        double *P = &S[(m + 3) % 3][0][0];
#pragma unroll
        for (int e = 0; e < EF_STAGE; e++) {
            const int p = t + e * EF_THREADS;
            if (soff[e] >= 0) {
                P[p] = pf[e][0];
                P[EF_PLANE + p] = pf[e][1];
                P[2 * EF_PLANE + p] = pf[e][2];
                P[3 * EF_PLANE + p] = pf[e][3];
            }
        }
    };
    // z windows of this thread's column: k-2 and k+2 in registers, k-1 .. k+1
    // from the staged planes
    const int ic = i < d.NX - 1 ? i : d.NX - 2, jc = j < d.NY - 1 ? j : d.NY - 2;
    const long col = (long)jc * d.px + ic;
    double zm2[4], zp2[4];
    {
        const long a2 = (long)(k0 - 2) * d.plane + col, p2 = a2 + 4 * d.plane;
        zm2[0] = A.u[a2]; zm2[1] = A.v[a2]; zm2[2] = A.w[a2]; zm2[3] = A.c[a2];
        zp2[0] = A.u[p2]; zp2[1] = A.v[p2]; zp2[2] = A.w[p2]; zp2[3] = A.c[p2];
    }
    stage(k0 - 1);
    stage(k0);
    stage(k0 + 1);
    if (k0 + 2 < d.NZ) fetch(k0 + 2);
    const int o = (ty + EF_H) * EF_SX + tx + EF_H;   // this thread's cell in a staged plane
    const double hx = A.h[0], hy = A.h[1], hz = A.h[2], rx = A.rh[0], ry = A.rh[1], rz = A.rh[2];
    for (int k = k0; k < k1; k++) {
        __syncthreads();
        if (act) {
            const double *Ubk = &S[(k + 2) % 3][0][0], *Vbk = &S[(k + 2) % 3][1][0];
            const double *Uk = &S[k % 3][0][0], *Vk = &S[k % 3][1][0], *Wk = &S[k % 3][2][0],
                         *Ck = &S[k % 3][3][0];
            const double *Uf = &S[(k + 1) % 3][0][0], *Vf = &S[(k + 1) % 3][1][0], *Wf = &S[(k + 1) % 3][2][0],
                         *Cf = &S[(k + 1) % 3][3][0];
            const double *Wbk = &S[(k + 2) % 3][2][0], *Cbk = &S[(k + 2) % 3][3][0];
            const long id = gidx(d, i, j, k);
            const double *zk = ZK[k - k0];
            // BOUNDARY statuses mark the box planes only, so none lies in the
            // fused region: every node here takes its convective term, and
            // the statuses and pressure gradients (read for the right-hand
            // sides and the particle velocity) are loaded first, their
            // latency overlapping the ENO arithmetic instead of gating it
            const int s_u = A.su[id], s_v = A.sv[id], s_w = A.sw[id], s_c = A.sc[id];
            double gx = 0.0, gy = 0.0, gz = 0.0;
            if constexpr (RHS) { gx = A.dpx[id]; gy = A.dpy[id]; gz = A.dpz[id]; }
            const double uc = Uk[o], vc = Vk[o], wc = Wk[o], cc = Ck[o];
            // u node
            {
                const double vu = interp4w(Vk[o - 1], vc, Vk[o + EF_SX - 1], Vk[o + EF_SX], wvu_x, wvu_y);
                const double wu = interp4w(Wk[o - 1], wc, Wf[o - 1], Wf[o], wwu_x, zk[0]);
                double r = eno_u5(Uk[o - 2], Uk[o - 1], uc, Uk[o + 1], Uk[o + 2], rx, hx, uc);
                r += eno_u5(Uk[o - 2 * EF_SX], Uk[o - EF_SX], uc, Uk[o + EF_SX], Uk[o + 2 * EF_SX], ry, hy, vu);
                r += eno_u5(zm2[0], Ubk[o], uc, Uf[o], zp2[0], rz, hz, wu);
                if (!RHS || A.store_conv) st_stream(&A.cu[id], r);
                if constexpr (RHS) st_stream(&A.ru[id], vel_row(s_u) ? uc * A.a_dt - r - gx : 0.0);
            }
            // v node
            {
                const double uv = interp4w(Uk[o - EF_SX], Uk[o - EF_SX + 1], uc, Uk[o + 1], wuv_x, wuv_y);
                const double wv = interp4w(Wk[o - EF_SX], Wf[o - EF_SX], wc, Wf[o], zk[1], wwv_y);
                double r = eno_u5(Vk[o - 2], Vk[o - 1], vc, Vk[o + 1], Vk[o + 2], rx, hx, uv);
                r += eno_u5(Vk[o - 2 * EF_SX], Vk[o - EF_SX], vc, Vk[o + EF_SX], Vk[o + 2 * EF_SX], ry, hy, vc);
                r += eno_u5(zm2[1], Vbk[o], vc, Vf[o], zp2[1], rz, hz, wv);
                if (!RHS || A.store_conv) st_stream(&A.cv[id], r);
                if constexpr (RHS) {
                    double rh = 0.0;
                    if (vel_row(s_v)) {
                        const double c_term = interp2d(Ck[o - EF_SX], ycm, cc, yc0, yv0);
                        rh = vc * A.a_dt - r - gy - A.buoy * c_term;
                    }
                    st_stream(&A.rv[id], rh);
                }
            }
            // w node
            {
                const double uw = interp4w(Ubk[o], Ubk[o + 1], uc, Uk[o + 1], wuw_x, zk[2]);
                const double vw = interp4w(Vbk[o], vc, Vbk[o + EF_SX], Vk[o + EF_SX], zk[3], wvw_y);
                double r = eno_u5(Wk[o - 2], Wk[o - 1], wc, Wk[o + 1], Wk[o + 2], rx, hx, uw);
                r += eno_u5(Wk[o - 2 * EF_SX], Wk[o - EF_SX], wc, Wk[o + EF_SX], Wk[o + 2 * EF_SX], ry, hy, vw);
                r += eno_u5(zm2[2], Wbk[o], wc, Wf[o], zp2[2], rz, hz, wc);
                if (!RHS || A.store_conv) st_stream(&A.cw[id], r);
                if constexpr (RHS) st_stream(&A.rw[id], vel_row(s_w) ? wc * A.a_dt - r - gz : 0.0);
            }
            // concentration cell
            {
                // Velocity_cell_center (Velocity.c:1476-1602) at this cell
                const double ubc = interp2d(uc, xu0, Uk[o + 1], xu1, xc0);
                double vbc = interp2d(vc, yv0, Vk[o + EF_SX], yv1, yc0);
                const double wbc = interp2d(wc, zk[4], Wf[o], zk[6 + 4], zk[5]);   // zw[k], zw[k+1], zc[k]
                if (A.particle) vbc = particle_v(vbc, s_c, cc, A.hindered, A.phi_max, A.v_settle);
                double r = eno_u5(Ck[o - 2], Ck[o - 1], cc, Ck[o + 1], Ck[o + 2], rx, hx, ubc);
                r += eno_u5(Ck[o - 2 * EF_SX], Ck[o - EF_SX], cc, Ck[o + EF_SX], Ck[o + 2 * EF_SX], ry, hy, vbc);
                r += eno_u5(zm2[3], Cbk[o], cc, Cf[o], zp2[3], rz, hz, wbc);
                if (!RHS || A.store_conv) st_stream(&A.cc[id], r);
                if constexpr (RHS) st_stream(&A.rc[id], s_c == TB_FLUID ? cc * A.a_dt - r : 0.0);
            }
            // the column's z window moves up: k-1 becomes k-2, k+3 arrives
            zm2[0] = Ubk[o]; zm2[1] = Vbk[o]; zm2[2] = Wbk[o]; zm2[3] = Cbk[o];
        }
        __syncthreads();
        // the slot of plane k - 1 is free: plane k + 2 (fetched in the
        // previous iteration) goes there, and plane k + 3 is fetched
        if (k + 2 < d.NZ) put(k + 2);
        if (k + 3 < d.NZ && k + 1 < k1) fetch(k + 3);
        if (act && k + 1 < k1) {
            const long p3 = (long)(k + 3) * d.plane + col;
            zp2[0] = A.u[p3]; zp2[1] = A.v[p3]; zp2[2] = A.w[p3]; zp2[3] = A.c[p3];
        }
    }
}

// the nodes within EF_SHELL of a box face (the fused kernel leaves them):
// each with the per-point code, over the x slabs, then the y slabs between
// them, then the z slabs between those
template <bool RHS>
__global__ void __launch_bounds__(256) k_eno_shell(MomArgs a, ConcArgs ca, long nx_slab, long ny_slab, long nz_slab) {
    const Dims &d = a.d;
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int H = EF_SHELL;
    int i, j, k;
    if (e < nx_slab) {
        const int side = (int)(e % (2 * H));
        i = side < H ? side : d.NX - 2 * H + side;
        const long r = e / (2 * H);
        j = (int)(r % d.NY);
        k = (int)(r / d.NY);
    } else if ((e -= nx_slab) < ny_slab) {
        const int w = d.NX - 2 * H;
        i = H + (int)(e % w);
        const long r = e / w;
        const int side = (int)(r % (2 * H));
        j = side < H ? side : d.NY - 2 * H + side;
        k = (int)(r / (2 * H));
    } else if ((e -= ny_slab) < nz_slab) {
        const int w = d.NX - 2 * H, h = d.NY - 2 * H;
        i = H + (int)(e % w);
        const long r = e / w;
        j = H + (int)(r % h);
        const int side = (int)(r / h);
        k = side < H ? side : d.NZ - 2 * H + side;
    } else {
        return;
    }
    mom_point<2, RHS>(a, i, j, k);
    conc_point<2, RHS>(ca, i, j, k);
}


// rhs_dt > 0: the four right-hand sides of step dt in the same pass, the
// convective terms stored only if store_conv
static MomArgs mom_args(tb_ctx *c, double rhs_dt, bool store_conv) {
    MomArgs a;
    a.d = c->d; a.g = c->g;
    a.W = c->eno_w;
    a.Wmax = c->eno_wmax;
    a.u = c->f[TB_F_U]; a.v = c->f[TB_F_V]; a.w = c->f[TB_F_W];
    a.su = c->st[0]; a.sv = c->st[1]; a.sw = c->st[2];
    a.cu = c->f[TB_F_CONV_U]; a.cv = c->f[TB_F_CONV_V]; a.cw = c->f[TB_F_CONV_W];
    for (int q = 0; q < 3; q++) a.L[q] = make_lines(c, q);
    a.dpx = c->f[TB_F_DPDX]; a.dpy = c->f[TB_F_DPDY]; a.dpz = c->f[TB_F_DPDZ];
    a.cfld = tb_buoyancy_source(c);
    a.ru = c->f[TB_F_RHS_U]; a.rv = c->f[TB_F_RHS_V]; a.rw = c->f[TB_F_RHS_W];
    a.a_dt = rhs_dt > 0.0 ? 1.0 / rhs_dt : 0.0;
    a.buoy = c->cfg.buoyancy;
    a.store_conv = store_conv ? 1 : 0;
    a.uin = tb_in_plane(c, TB_Q_U);
    a.uout = tb_out_plane(c, TB_Q_U);
    return a;
}

static int eno_momentum(tb_ctx *c, double rhs_dt, bool store_conv) {
    if (eno_weights(c)) return -1;
    const MomArgs a = mom_args(c, rhs_dt, store_conv);
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    const dim3 grid = tiles_grid(t), blk = tiles_block();
    const bool rhs = rhs_dt > 0.0;
    if (c->prof_on) tb_prof_begin(c, TB_K_ENO_MOM);
    switch (c->cfg.eno_order) {
    case 1:
        if (rhs) hipLaunchKernelGGL((k_eno_momentum<1, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_momentum<1, false>), grid, blk, 0, c->s, a, t);
        break;
    case 2:
        if (rhs) hipLaunchKernelGGL((k_eno_momentum<2, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_momentum<2, false>), grid, blk, 0, c->s, a, t);
        break;
    default:
        if (rhs) hipLaunchKernelGGL((k_eno_momentum<3, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_momentum<3, false>), grid, blk, 0, c->s, a, t);
        break;
    }
    if (c->prof_on) tb_prof_end(c, TB_K_ENO_MOM);
    TB_LAUNCH_CHECK(c);
    return 0;
}

int tb_eno_momentum(tb_ctx *c) { return eno_momentum(c, 0.0, true); }

static ConcArgs conc_args(tb_ctx *c, double rhs_dt, bool store_conv) {
    ConcArgs a;
    a.d = c->d; a.g = c->g;
    a.u = c->f[TB_F_U]; a.v = c->f[TB_F_V]; a.w = c->f[TB_F_W]; a.c = c->f[TB_F_C];
    a.sc = c->st[3];
    a.cc = c->f[TB_F_CONV_C];
    a.v_settle = c->cfg.v_settle;
    a.particle = (c->cfg.conc_type == TB_PARTICLE);
    a.hindered = a.particle && c->hindered;
    a.phi_max = c->cs[c->conc_sel].p.phi_max;
    a.L = make_lines(c, TB_Q_C);
    a.rc = c->f[TB_F_RHS_C];
    a.a_dt = rhs_dt > 0.0 ? 1.0 / rhs_dt : 0.0;
    a.store_conv = store_conv ? 1 : 0;
    a.uin = tb_in_plane(c, TB_Q_U);
    a.uout = tb_out_plane(c, TB_Q_U);
    return a;
}

static int eno_conc(tb_ctx *c, double rhs_dt, bool store_conv) {
    const ConcArgs a = conc_args(c, rhs_dt, store_conv);
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    const dim3 grid = tiles_grid(t), blk = tiles_block();
    const bool rhs = rhs_dt > 0.0;
    if (c->prof_on) tb_prof_begin(c, TB_K_ENO_CONC);
    switch (c->cfg.eno_order) {
    case 1:
        if (rhs) hipLaunchKernelGGL((k_eno_conc<1, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_conc<1, false>), grid, blk, 0, c->s, a, t);
        break;
    case 2:
        if (rhs) hipLaunchKernelGGL((k_eno_conc<2, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_conc<2, false>), grid, blk, 0, c->s, a, t);
        break;
    default:
        if (rhs) hipLaunchKernelGGL((k_eno_conc<3, true>), grid, blk, 0, c->s, a, t);
        else hipLaunchKernelGGL((k_eno_conc<3, false>), grid, blk, 0, c->s, a, t);
        break;
    }
    if (c->prof_on) tb_prof_end(c, TB_K_ENO_CONC);
    TB_LAUNCH_CHECK(c);
    return 0;
}

int tb_eno_conc(tb_ctx *c) { return eno_conc(c, 0.0, true); }

// ENO of all four quantities and their right-hand sides in one pass per
// kernel: the same values as tb_eno_conv(c, -1) followed by tb_set_rhs for
// q = 0..3 (the right-hand-side formulas of k_rhs_vel / k_rhs_conc applied to
// the convective term in registers), without re-reading the convective terms
// and fields.  tb_step passes store_conv = false (nothing downstream reads
// the convective terms: the TB_F_CONV_* fields keep their last values).
// Second order: u, v, w and c in one k-marching pass (k_eno_fused2) plus
// the zeros of the last planes
static int eno_fused(tb_ctx *c, double rhs_dt, bool store_conv) {
    if (eno_weights(c)) return -1;
    const MomArgs m = mom_args(c, rhs_dt, store_conv);
    const ConcArgs ca = conc_args(c, rhs_dt, store_conv);
    FusedArgs A;
    A.d = c->d;
    A.u = m.u; A.v = m.v; A.w = m.w; A.c = ca.c;
    A.su = m.su; A.sv = m.sv; A.sw = m.sw; A.sc = ca.sc;
    A.dpx = m.dpx; A.dpy = m.dpy; A.dpz = m.dpz;
    A.cu = m.cu; A.cv = m.cv; A.cw = m.cw; A.cc = ca.cc;
    A.ru = m.ru; A.rv = m.rv; A.rw = m.rw; A.rc = ca.rc;
    A.W = m.W;
    A.wm = m.Wmax;
    A.xu = c->g.xu; A.xc = c->g.xc; A.yv = c->g.yv; A.yc = c->g.yc; A.zw = c->g.zw; A.zc = c->g.zc;
    A.a_dt = m.a_dt;
    A.buoy = m.buoy;
    A.v_settle = ca.v_settle;
    A.particle = ca.particle;
    A.hindered = ca.hindered;
    A.phi_max = ca.phi_max;
    A.store_conv = m.store_conv;
    // the fast path needs one spacing per axis: a quantity whose uniform
    // spacing differs from the x / y / z one of u, v, w, c's first uniform
    // line takes the general form
    A.unimask = 0;
    for (int ax = 0; ax < 3; ax++) {
        A.h[ax] = 0.0;
        A.rh[ax] = 0.0;
        A.hi[ax] = ax == 0 ? c->d.NX - 2 : (ax == 1 ? c->d.NY - 2 : c->d.NZ - 2);
        for (int q = 0; q < 4; q++) {
            const LineSpec L = (q < 3 ? m.L[q] : ca.L).ax[ax];
            if (!L.uni || L.lo != (q == ax ? 1 : 0) || L.hi != A.hi[ax]) continue;
            if (A.h[ax] == 0.0) { A.h[ax] = L.h; A.rh[ax] = L.rh; }
            if (L.h == A.h[ax] && L.rh == A.rh[ax]) A.unimask |= 1 << (3 * q + ax);
        }
    }
    const Dims &d = c->d;
    const bool rhs = rhs_dt > 0.0;
    // several fields: the v buoyancy reads c_total, not the staged field
    if (c->nconc > 1) return 1;
    if (A.unimask != 0xFFF || d.NX < 2 * EF_SHELL + 2 || d.NY < 2 * EF_SHELL + 2 || d.NZ < 2 * EF_SHELL + 2)
        return 1;   // not this kernel's grid: the caller runs the per-point kernels
    const EfTiles T = make_ef_tiles(d, EF_SHELL);
    const dim3 grid((unsigned)(T.per_xcd * NXCD)), blk(EF_TX, EF_TY);
    if (c->prof_on) tb_prof_begin(c, TB_K_ENO_MOM);
    if (rhs) hipLaunchKernelGGL(k_eno_fused2<true>, grid, blk, 0, c->s, A, T);
    else hipLaunchKernelGGL(k_eno_fused2<false>, grid, blk, 0, c->s, A, T);
    if (c->prof_on) tb_prof_end(c, TB_K_ENO_MOM);
    const long H = EF_SHELL;
    const long nxs = 2 * H * d.NY * d.NZ, nys = (d.NX - 2 * H) * 2 * H * d.NZ,
               nzs = (d.NX - 2 * H) * (d.NY - 2 * H) * 2 * H;
    const long np = nxs + nys + nzs;
    if (c->prof_on) tb_prof_begin(c, TB_K_ENO_CONC);
    if (rhs)
        hipLaunchKernelGGL(k_eno_shell<true>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->s, m, ca, nxs,
                           nys, nzs);
    else
        hipLaunchKernelGGL(k_eno_shell<false>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->s, m, ca, nxs,
                           nys, nzs);
    if (c->prof_on) tb_prof_end(c, TB_K_ENO_CONC);
    TB_LAUNCH_CHECK(c);
    return 0;
}

// every concentration field is computed (each Conc_* call of the reference
// takes its c[iconc])
static int eno_rhs_kernels(tb_ctx *c, double dt, bool store_conv) {
    if (c->cfg.eno_order == 2) {
        const int rc = eno_fused(c, dt, store_conv);
        if (rc <= 0) return rc;
    }
    if (eno_momentum(c, dt, store_conv)) return -1;
    if (c->nconc == 1) return eno_conc(c, dt, store_conv);
    const int sel = c->conc_sel;
    for (int n = 0; n < c->nconc; n++)
        if (tb_conc_use(c, n) || eno_conc(c, dt, store_conv)) return -1;
    return tb_conc_use(c, sel);
}

// and the open faces' right-hand-side terms after them
int tb_eno_rhs_async(tb_ctx *c, double dt, bool store_conv) {
    if (eno_rhs_kernels(c, dt, store_conv)) return -1;
    return tb_open_rhs(c, -1);
}

extern "C" int tb_eno_conv_rhs(tb_ctx *c, double dt) {
    if (!c || !(dt > 0.0)) { if (c) c->err = "tb_eno_conv_rhs: bad argument"; return -1; }
    if (tb_eno_rhs_async(c, dt, true)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_eno_conv(tb_ctx *c, int q) {
    if (!c || q < -1 || q > 3) { if (c) c->err = "tb_eno_conv: bad quantity"; return -1; }
    int rc = 0;
    if (q == -1 && c->cfg.eno_order == 2) rc = eno_fused(c, 0.0, true);
    if (rc < 0) return rc;
    if (!(q == -1 && c->cfg.eno_order == 2) || rc > 0) {
        rc = 0;
        if (q != TB_Q_C) rc = tb_eno_momentum(c);   // u, v, w are produced together
        if (rc == 0 && (q == TB_Q_C || q == -1)) rc = tb_eno_conc(c);
    }
    if (rc == 0) TB_CHECK(c, hipStreamSynchronize(c->s));
    return rc;
}

// ---------------------------------------------------------------------------
// Right-hand sides: Velocity_u/v/w_set_RHS (Velocity.c:1607-2043),
// Conc_set_RHS (Conc.c:864-992).

__global__ void k_rhs_vel(Dims d, GridDev g, int q, const double *__restrict__ data, const double *__restrict__ conv,
                          const double *__restrict__ dp, const double *__restrict__ cfld,
                          const uint8_t *__restrict__ st, double *__restrict__ rhs, double a_dt, double buoy,
                          Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    int s = st[id];
    double r = 0.0;
    if (s == TB_FLUID || s == TB_BUFFER_SOLID) {
        if (q == TB_Q_V) {
            double c_term = interp2d(cfld[id - d.px], g.yc[j - 1], cfld[id], g.yc[j], g.yv[j]);
            r = data[id] * a_dt - conv[id] - dp[id] - buoy * c_term;
        } else {
            r = data[id] * a_dt - conv[id] - dp[id];
        }
    }
    rhs[id] = r;
}

__global__ void k_rhs_conc(Dims d, const double *__restrict__ data, const double *__restrict__ conv,
                           const uint8_t *__restrict__ st, double *__restrict__ rhs, double a_dt, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    rhs[id] = (st[id] == TB_FLUID) ? data[id] * a_dt - conv[id] : 0.0;
}

extern "C" int tb_set_rhs(tb_ctx *c, int q, double dt) {
    if (!c || q < 0 || q > 3) { if (c) c->err = "tb_set_rhs: bad quantity"; return -1; }
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    double a_dt = 1.0 / dt;
    if (q == TB_Q_C)
        hipLaunchKernelGGL(k_rhs_conc, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->f[TB_F_C], c->f[TB_F_CONV_C],
                           c->st[3], c->f[TB_F_RHS_C], a_dt, t);
    else
        hipLaunchKernelGGL(k_rhs_vel, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->g, q, c->f[TB_F_U + q],
                           c->f[TB_F_CONV_U + q], c->f[TB_F_DPDX + q], tb_buoyancy_source(c), c->st[q],
                           c->f[TB_F_RHS_U + q],
                           a_dt, c->cfg.buoyancy, t);
    TB_LAUNCH_CHECK(c);
    return tb_open_rhs(c, q);
}

// ---------------------------------------------------------------------------
// Extract_total_kinetic_energy (Extract.c:803-884) and
// Conc_compute_total_suspended_mass (Conc.c:1471-1512) in one pass over the
// cells: over FLUID cells (concentration status), the sums of
// (ubc^2 + vbc^2 + wbc^2) dV with the cell-centre velocities of
// Velocity_cell_center (Velocity.c:1476-1602) and dV = (dx dy) dz, and of
// c dV with dV = (dz dy) dx, each as the reference associates it.
__global__ void __launch_bounds__(256) k_diag(Dims d, GridDev g, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ w,
                                              const double *__restrict__ c, const uint8_t *__restrict__ sc,
                                              double *part, Tiles t, const double *__restrict__ uin,
                                              const double *__restrict__ uout) {
    __shared__ double sh[48];
    double acc[3] = {0.0, 0.0, 0.0};
    int i, j, k;
    if (tile_point(t, i, j, k)) {
        const long id = gidx(d, i, j, k);
        if (sc[id] == TB_FLUID) {
            const long pidx = (long)k * d.NY + j;
            const double u0 = (uin && i == 0) ? uin[pidx] : u[id];
            const double u1 = (uout && i == d.NX - 2) ? uout[pidx] : u[id + 1];
            const double ubc = interp2d(u0, g.xu[i], u1, g.xu[i + 1], g.xc[i]);
            const double vbc = interp2d(v[id], g.yv[j], v[id + d.px], g.yv[j + 1], g.yc[j]);
            const double wbc = interp2d(w[id], g.zw[k], w[id + d.plane], g.zw[k + 1], g.zc[k]);
            const double dx = g.dxc[i], dy = g.dyc[j], dz = g.dzc[k];
            acc[0] = (ubc * ubc + vbc * vbc + wbc * wbc) * (dx * dy * dz);
            acc[1] = c[id] * (dz * dy * dx);
            acc[2] = c[id] * (dx * dy * dz);   // Extract_Stokes_dissipation_rate's dV
        }
    }
    block_sum<3>(acc, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[blockIdx.x] = acc[0];
        part[RED_MAX_BLOCKS + blockIdx.x] = acc[1];
        part[2 * RED_MAX_BLOCKS + blockIdx.x] = acc[2];
    }
}

static int diag_sums(tb_ctx *c, double out[3]);

static int diag_sums(tb_ctx *c, double out[3]) {
    const Tiles t = make_tiles(c->ncx, c->ncy, c->ncz);
    const dim3 grid = tiles_grid(t);
    if ((long)grid.x * grid.y * grid.z > RED_MAX_BLOCKS || grid.y != 1 || grid.z != 1) {
        c->err = "tb_diagnostics: grid too large for the reduction buffer";
        return -1;
    }
    hipLaunchKernelGGL(k_diag, grid, tiles_block(), 0, c->s, c->d, c->g, (const double *)c->f[TB_F_U],
                       (const double *)c->f[TB_F_V], (const double *)c->f[TB_F_W], (const double *)c->f[TB_F_C],
                       (const uint8_t *)c->st[3], c->red, t, tb_in_plane(c, TB_Q_U), tb_out_plane(c, TB_Q_U));
    TB_LAUNCH_CHECK(c);
    if (tb_reduce_finish(c, (int)grid.x, 3, S_TMP0)) return -1;
    if (tb_read_scalars(c, S_TMP0, 3)) return -1;
    for (int a = 0; a < 3; a++) out[a] = c->hscal[S_TMP0 + a];
    return 0;
}

extern "C" int tb_diagnostics(tb_ctx *c, double *kinetic_energy, double *suspended_mass) {
    if (!c) return -1;
    double s[3];
    if (diag_sums(c, s)) return -1;
    if (kinetic_energy) *kinetic_energy = 0.5 * s[0];
    if (suspended_mass) *suspended_mass = s[1];
    return 0;
}

// Extract_Stokes_dissipation_rate (Extract.c:886-950): |v_settl0| alpha
// times the sum of c dV over FLUID cells for a particle concentration, 0
// for salinity
extern "C" int tb_stokes_dissipation(tb_ctx *c, double conc_alpha, double *rate) {
    if (!c || !rate) {
        if (c) c->err = "tb_stokes_dissipation: bad argument";
        return -1;
    }
    if (c->cfg.conc_type != TB_PARTICLE) {
        *rate = 0.0;
        return 0;
    }
    double s[3];
    if (diag_sums(c, s)) return -1;
    *rate = s[2] * (fabs(c->cfg.v_settle) * conc_alpha);
    return 0;
}

// ---------------------------------------------------------------------------
// Conc_bottom_sed_rate (Conc.c:2158-2229) with the interface index of
// Grid_set_interface_y_index (Grid.c:1972-2009): per (i, k) column the first
// IMMERSED cell from the bottom, or row 0 when it is FLUID; the column's
// sedimentation rate u_s c dx dz there, 0 where no such cell exists.
__global__ void k_sed_bottom(Dims d, GridDev g, const double *__restrict__ c, const uint8_t *__restrict__ sc,
                             double us, double *__restrict__ out) {
    const long n = (long)d.NX * d.NZ;
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    int jf = -1;
    for (int j = 0; j < d.NY; j++) {
        const int s = sc[gidx(d, i, j, k)];
        if (s == TB_IMMERSED || (j == 0 && s == TB_FLUID)) {
            jf = j;
            break;
        }
    }
    out[a] = jf >= 0 ? us * c[gidx(d, i, jf, k)] * g.dxc[i] * g.dzc[k] : 0.0;
}

// Conc_compute_sedimentation_rate (Conc.c:2098-2157): the column values
// summed in the reference's order (k outer, i inner) on the host, so the
// total is the reference's bits; bottom (NZ x NX, x fastest) may be NULL.
extern "C" int tb_sedimentation_rate(tb_ctx *c, double *total, double *bottom) {
    if (!c) return -1;
    const Dims &d = c->d;
    const long n = (long)d.NX * d.NZ;
    std::vector<double> h((size_t)n, 0.0);
    if (c->cfg.conc_type == TB_PARTICLE) {
        double *dev = c->kv[7];   // scratch vector (any field-sized buffer of the context)
        hipLaunchKernelGGL(k_sed_bottom, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, d, c->g,
                           (const double *)c->f[TB_F_C], (const uint8_t *)c->st[3], fabs(c->cfg.v_settle), dev);
        TB_LAUNCH_CHECK(c);
        TB_CHECK(c, hipMemcpyAsync(h.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
        TB_CHECK(c, hipStreamSynchronize(c->s));
    }
    double s = 0.0;
    for (long a = 0; a < n; a++) s += h[a];
    if (total) *total = s;
    if (bottom) memcpy(bottom, h.data(), sizeof(double) * n);
    return 0;
}

// ---------------------------------------------------------------------------
// CFL time step.  The reference keeps its parameters (cfl_method, alpha = 0.5,
// default_dt, Input.c:421-423) but its rule lived in the unpublished driver;
// the rule here: over the FLUID cells of the concentration grid, with the
// cell-centre velocities of Velocity_cell_center (Velocity.c:1476-1602),
// rate = |u_c|/dx + |v_c|/dy + |w_c|/dz, and dt = min(dt_max, alpha / max rate).
// Per-workgroup maxima, then one workgroup takes the maximum of those (a
// maximum does not depend on the order, so the result is exact); a NaN
// anywhere wins every comparison and is reported.
__device__ inline double max_nan(double a, double b) { return (a != a || a > b) ? a : b; }

__global__ void __launch_bounds__(256) k_cfl(Dims d, GridDev g, const double *__restrict__ u,
                                             const double *__restrict__ v, const double *__restrict__ w,
                                             const uint8_t *__restrict__ sc, double *part, Tiles t,
                                             const double *__restrict__ uin, const double *__restrict__ uout) {
    __shared__ double sh[256];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double r = 0.0;
    int i, j, k;
    if (tile_point(t, i, j, k)) {
        const long id = gidx(d, i, j, k);
        if (sc[id] == TB_FLUID) {
            const long pidx = (long)k * d.NY + j;
            const double u0 = (uin && i == 0) ? uin[pidx] : u[id];
            const double u1 = (uout && i == d.NX - 2) ? uout[pidx] : u[id + 1];
            const double ubc = interp2d(u0, g.xu[i], u1, g.xu[i + 1], g.xc[i]);
            const double vbc = interp2d(v[id], g.yv[j], v[id + d.px], g.yv[j + 1], g.yc[j]);
            const double wbc = interp2d(w[id], g.zw[k], w[id + d.plane], g.zw[k + 1], g.zc[k]);
            r = fabs(ubc) / g.dxc[i] + fabs(vbc) / g.dyc[j] + fabs(wbc) / g.dzc[k];
        }
    }
    sh[tid] = r;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = max_nan(sh[tid], sh[tid + s]);
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_max_final(const double *part, long n, double *out) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (long b = threadIdx.x; b < n; b += blockDim.x) m = max_nan(m, part[b]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = max_nan(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

extern "C" int tb_cfl_dt(tb_ctx *c, double alpha, double dt_max, double *dt, double *rate) {
    if (!c || !dt || !(alpha > 0.0) || !(dt_max > 0.0)) {
        if (c) c->err = "tb_cfl_dt: bad argument";
        return -1;
    }
    const Tiles t = make_tiles(c->ncx, c->ncy, c->ncz);
    const dim3 grid = tiles_grid(t);
    if ((long)grid.x > RED_MAX_BLOCKS) {
        c->err = "tb_cfl_dt: grid too large for the reduction buffer";
        return -1;
    }
    hipLaunchKernelGGL(k_cfl, grid, tiles_block(), 0, c->s, c->d, c->g, (const double *)c->f[TB_F_U],
                       (const double *)c->f[TB_F_V], (const double *)c->f[TB_F_W], (const uint8_t *)c->st[3], c->red,
                       t, tb_in_plane(c, TB_Q_U), tb_out_plane(c, TB_Q_U));
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(1024), 0, c->s, (const double *)c->red, (long)grid.x,
                       c->scal + S_TMP0);
    TB_LAUNCH_CHECK(c);
    if (tb_read_scalars(c, S_TMP0, 1)) return -1;
    const double m = c->hscal[S_TMP0];
    if (!(m <= DBL_MAX)) {
        c->err = "tb_cfl_dt: NaN or Inf velocity";
        return -1;
    }
    if (rate) *rate = m;
    *dt = (m > 0.0 && alpha / m < dt_max) ? alpha / m : dt_max;
    return 0;
}
