// tb_poisson.hip -- pressure Poisson right-hand side, operator, geometric
// multigrid (FP64 and FP32), GMG-preconditioned CG, mixed-precision
// iterative refinement, projection and the divergence check.
//
// The reference assembles a 7-point -Laplacian over the whole box with
// Neumann walls and one pinned corner cell (Pressure_form_LHS_matrix,
// Pressure.c:691-936) and hands it to GMRES + BoomerAMG.  Here the operator
// is applied matrix-free with the same row values; the solve runs CG on the
// Neumann operator (the same matrix without the pin) with a symmetric V-cycle
// preconditioner, then shifts the result so the pinned cell holds 0.  For a
// compatible right-hand side (sum b = 0, which the divergence RHS satisfies
// up to rounding) that is the unique solution of the pinned system.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "tb_common.h"

// ---------------------------------------------------------------------------
// Pressure_set_RHS (Pressure.c:171-276)
// u on an open x face from its plane (Pressure.c:242-249; null: closed)
__global__ void k_prhs(Dims d, GridDev g, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ w, const uint8_t *__restrict__ sc, double *__restrict__ rhs,
                       double a_dEta, double a_dXi, double a_dPhi, double a_dt, Tiles t,
                       const double *__restrict__ uin, const double *__restrict__ uout) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    double r = 0.0;
    if (sc[id] != TB_BOUNDARY) {
        double C_z = g.mzc[k] * a_dPhi;
        double C_y = g.myc[j] * a_dXi;
        double C_x = g.mxc[i] * a_dEta;
        double dudx = (u[id + 1] - u[id]) * C_x;
        if (uin && i == 0) dudx = (u[id + 1] - uin[(long)k * d.NY + j]) * C_x;
        if (uout && i == d.NX - 2) dudx = (uout[(long)k * d.NY + j] - u[id]) * C_x;
        double dvdy = (v[id + d.px] - v[id]) * C_y;
        double dwdz = (w[id + d.plane] - w[id]) * C_z;
        r = -(dudx + dvdy + dwdz) * a_dt;
    }
    rhs[id] = r;
}

extern "C" int tb_poisson_rhs(tb_ctx *c, double dt) {
    if (!c) return -1;
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    hipLaunchKernelGGL(k_prhs, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->g, c->f[TB_F_U], c->f[TB_F_V],
                       c->f[TB_F_W], c->st[3], c->f[TB_F_PRHS], 1.0 / c->dEta, 1.0 / c->dXi, 1.0 / c->dPhi, 1.0 / dt, t,
                       tb_in_plane(c, TB_Q_U), tb_out_plane(c, TB_Q_U));
    TB_LAUNCH_CHECK(c);
    return 0;
}

// Pressure_compute_real_pressure (Pressure.c:590-685): b = -div(grad p) from
// the stored pressure gradients (0 on BOUNDARY cells), the expression of
// k_prhs with the gradients for the velocities and a factor 1.0 (exact) for
// 1/dt; tb_real_pressure then solves into phi, which (as in the reference)
// is the next solve's initial guess.
extern "C" int tb_real_pressure_rhs(tb_ctx *c) {
    if (!c) return -1;
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    hipLaunchKernelGGL(k_prhs, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->g, c->f[TB_F_DPDX], c->f[TB_F_DPDY],
                       c->f[TB_F_DPDZ], c->st[3], c->f[TB_F_PRHS], 1.0 / c->dEta, 1.0 / c->dXi, 1.0 / c->dPhi, 1.0,
                       t, (const double *)nullptr, (const double *)nullptr);
    TB_LAUNCH_CHECK(c);
    return 0;
}

// ---------------------------------------------------------------------------
// Operator rows (Pressure.c:740-924); neumann = 1 drops the corner pin.
struct PoisOp {
    Dims d;
    GridDev g;
    double Cx, Cy, Cz, ap_solid;
    int neumann;
};

// k-marching application of the operator.  RESID = 0: y = A x with partial
// x.y ; RESID = 1: r = b - A x (and r32 = (float) r if given) with partials
// r.r and b.b.  x and y parts of the coefficients are formed once per column,
// z parts once per plane; k-1 / k / k+1 stay in registers.
template <int RESID>
__global__ void __launch_bounds__(256) k_pois_km(PoisOp P, const double *__restrict__ x, const double *__restrict__ b,
                                                 double *__restrict__ y, float *__restrict__ y32, double *part,
                                                 KTiles t) {
    __shared__ double sh[48];
    double acc[3] = {0.0, 0.0, 0.0};
    int i, j, k0, k1;
    bool act;
    ktile_column(t, i, j, k0, k1, act);
    if (act) {
        const Dims &d = P.d;
        const GridDev &g = P.g;
        const int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
        const double Cx = P.Cx, Cy = P.Cy, Cz = P.Cz;
        const bool bnd_col = (i == NI) || (j == NJ);
        const int ip1 = i + 1 <= NI ? i + 1 : NI, jp1 = j + 1 <= NJ ? j + 1 : NJ;
        const double mpx = g.mxc[i], mex = g.mxu[ip1], mwx = g.mxu[i];
        const double mpy = g.myc[j], mny = g.myv[jp1], msy = g.myv[j];
        const double axy = -Cx * mpx * (mex + mwx) - Cy * mpy * (mny + msy);
        const double aw = Cx * mpx * mwx, ae = Cx * mpx * mex, an = Cy * mpy * mny, as = Cy * mpy * msy;
        const bool hs = j > 0, hw = i > 0, he = i < NI - 1, hn = j < NJ - 1;
        const bool pin_col = P.neumann == 0 && i == NI - 1 && j == NJ - 1;
        // neighbour offsets clamped to the point itself when absent, so all
        // loads are issued unconditionally and overlap
        const long os = hs ? d.px : 0, on = (j < NJ) ? d.px : 0;
        const long ow = hw ? 1 : 0, oe = (i < NI) ? 1 : 0;
        long id = gidx(d, i, j, k0);
        double xm = x[id - ((k0 > 0) ? d.plane : 0)];
        double xc = x[id];
        for (int k = k0; k < k1; k++) {
            double xp = x[id + ((k + 1 <= NK) ? d.plane : 0)];
            const double xs = x[id - os], xw = x[id - ow], xe = x[id + oe], xn = x[id + on];
            const double bi = RESID ? b[id] : 0.0;
            double v;
            if (bnd_col || k == NK) {
                v = P.ap_solid * xc;   // BOUNDARY half-cell rows
            } else {
                const double mpz = g.mzc[k], mfz = g.mzw[k + 1], mbz = g.mzw[k];
                double ap = axy - Cz * mpz * (mbz + mfz);
                const double ab = Cz * mpz * mbz, af = Cz * mpz * mfz;
                const bool hb = k > 0, hf = k < NK - 1;
                if (!hb) ap += ab;
                if (!hs) ap += as;
                if (!hw) ap += aw;
                if (!he && !(pin_col && k == NK - 1)) ap += ae;
                if (!hn) ap += an;
                if (!hf) ap += af;
                double r = 0.0;
                r = hb ? r + ab * xm : r;
                r = hs ? r + as * xs : r;
                r = hw ? r + aw * xw : r;
                r += ap * xc;
                r = he ? r + ae * xe : r;
                r = hn ? r + an * xn : r;
                r = hf ? r + af * xp : r;
                v = r;
            }
            if (RESID) {
                double ri = bi - v;
                if (y) st_stream(&y[id], ri);
                if (y32) st_stream(&y32[id], (float)ri);
                acc[0] += ri * ri;
                acc[1] += bi * bi;
                acc[2] += xc * bi;
            } else {
                st_stream(&y[id], v);
                acc[0] += xc * v;
            }
            xm = xc;
            xc = xp;
            id += d.plane;
        }
    }
    block_sum<3>(acc, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[blockIdx.x] = acc[0];
        part[RED_MAX_BLOCKS + blockIdx.x] = acc[1];
        part[2 * RED_MAX_BLOCKS + blockIdx.x] = acc[2];
    }
}

static PoisOp make_pois(tb_ctx *c, int neumann) {
    PoisOp P;
    P.d = c->d; P.g = c->g;
    P.Cx = -1.0 / (c->dEta * c->dEta);
    P.Cy = -1.0 / (c->dXi * c->dXi);
    P.Cz = -1.0 / (c->dPhi * c->dPhi);
    P.ap_solid = -P.Cx * c->hgrid[15][0] - P.Cy * c->hgrid[16][0] - P.Cz * c->hgrid[17][0];
    P.neumann = neumann;
    return P;
}

extern "C" int tb_poisson_apply(tb_ctx *c, int f_in, int f_out) {
    if (!c || f_in < 0 || f_in >= TB_F_COUNT || f_out < 0 || f_out >= TB_F_COUNT || f_in == f_out) {
        if (c) c->err = "tb_poisson_apply: bad argument";
        return -1;
    }
    tb_field_written(c, f_out);
    KTiles t = make_ktiles(c->d.NX, c->d.NY, c->d.NZ);
    hipLaunchKernelGGL(k_pois_km<0>, ktiles_grid(t), tiles_block(), 0, c->s, make_pois(c, 0), (const double *)c->f[f_in],
                       (const double *)nullptr, c->f[f_out], (float *)nullptr, c->red, t);
    TB_LAUNCH_CHECK(c);
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

// ---------------------------------------------------------------------------
// Geometric multigrid on cell-centred levels.  Level 0 is the interior of
// the padded field layout; coarser levels are compact.  Coarsening halves a
// direction when its spacing is at most half the largest one (semi-coarsening
// the strongly coupled directions first; all directions when the spacings
// are within a factor 2), stops at <= COARSEST_CELLS cells,
// and the coarsest level is solved with a dense pseudo-inverse (A + J/n)^-1.

constexpr long COARSEST_CELLS = 64;

template <typename T>
struct LV {
    int nx, ny, nz;
    int px, py;          // row pitch, rows per plane
    T cx, cy, cz;        // 1/h^2
};

template <typename T>
__device__ inline long lidx(const LV<T> &L, int i, int j, int k) {
    return ((long)k * L.py + j) * L.px + i;
}

template <typename T>
__global__ void __launch_bounds__(256) k_zero(LV<T> L, T *__restrict__ x, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    x[lidx(L, i, j, k)] = (T)0;
}

// Per-cell bodies of the multigrid operations, shared by the one-launch-per-
// step kernels of the large levels and the single-workgroup kernel of the
// small ones (so both give the same bits).

// red-black Gauss-Seidel update of cell (i, j, k); zero_init: the other
// colour still holds the zero initial guess, so no neighbour is read
// the Gauss-Seidel value of cell (i, j, k) from b and its six neighbour
// values (values of absent neighbours are not used)
template <typename T>
__device__ inline T rb_value(const LV<T> &L, int i, int j, int k, T bv, T xw, T xe, T xs, T xn, T xb, T xf) {
    const bool hw = i > 0, he = i < L.nx - 1, hs = j > 0, hn = j < L.ny - 1, hb = k > 0, hf = k < L.nz - 1;
    T s = bv, dg = (T)0;
    s = hw ? s + L.cx * xw : s; dg = hw ? dg + L.cx : dg;
    s = he ? s + L.cx * xe : s; dg = he ? dg + L.cx : dg;
    s = hs ? s + L.cy * xs : s; dg = hs ? dg + L.cy : dg;
    s = hn ? s + L.cy * xn : s; dg = hn ? dg + L.cy : dg;
    s = hb ? s + L.cz * xb : s; dg = hb ? dg + L.cz : dg;
    s = hf ? s + L.cz * xf : s; dg = hf ? dg + L.cz : dg;
    // the FP32 smoother (inside the FP64 refinement) takes the hardware
    // reciprocal, within 1 ulp, one instruction instead of a correctly
    // rounded division's ten; the FP64 one divides exactly
    if constexpr (sizeof(T) == 4) return s * __builtin_amdgcn_rcpf(dg);
    else return s / dg;
}

// the Gauss-Seidel value of cell (i, j, k) from x and b (id: its index)
template <typename T>
__device__ inline T rbgs_value(const LV<T> &L, const T *x, const T *b, int i, int j, int k, bool zero_init,
                               long &id) {
    id = lidx(L, i, j, k);
    const long pl = (long)L.px * L.py;
    const bool hw = i > 0, he = i < L.nx - 1, hs = j > 0, hn = j < L.ny - 1, hb = k > 0, hf = k < L.nz - 1;
    T xw = (T)0, xe = (T)0, xs = (T)0, xn = (T)0, xb = (T)0, xf = (T)0;
    if (!zero_init) {
        // all neighbour loads issued up front (clamped to the cell itself)
        xw = x[id - (hw ? 1 : 0)]; xe = x[id + (he ? 1 : 0)];
        xs = x[id - (hs ? L.px : 0)]; xn = x[id + (hn ? L.px : 0)];
        xb = x[id - (hb ? pl : 0)]; xf = x[id + (hf ? pl : 0)];
    }
    return rb_value(L, i, j, k, b[id], xw, xe, xs, xn, xb, xf);
}

template <typename T>
__device__ inline void rbgs_cell(const LV<T> &L, T *__restrict__ x, const T *__restrict__ b, int i, int j, int k,
                                 bool zero_init) {
    long id;
    const T v = rbgs_value(L, (const T *)x, b, i, j, k, zero_init, id);
    x[id] = v;
}

// b - A x at cell (i, j, k) (id: its index)
template <typename T>
__device__ inline T resid_value(const LV<T> &L, const T *x, const T *b, int i, int j, int k, long &id) {
    id = lidx(L, i, j, k);
    const long pl = (long)L.px * L.py;
    const bool hw = i > 0, he = i < L.nx - 1, hs = j > 0, hn = j < L.ny - 1, hb = k > 0, hf = k < L.nz - 1;
    // absent neighbours are clamped to the cell itself, contributing xc - xc = 0
    const T xc = x[id];
    const T xw = x[id - (hw ? 1 : 0)], xe = x[id + (he ? 1 : 0)];
    const T xs = x[id - (hs ? L.px : 0)], xn = x[id + (hn ? L.px : 0)];
    const T xb = x[id - (hb ? pl : 0)], xf = x[id + (hf ? pl : 0)];
    const T bb = b[id];
    T s = L.cx * (xc - xw) + L.cx * (xc - xe) + L.cy * (xc - xs) + L.cy * (xc - xn) + L.cz * (xc - xb) +
          L.cz * (xc - xf);
    return bb - s;
}

template <typename T>
__device__ inline void resid_cell(const LV<T> &L, const T *__restrict__ x, const T *__restrict__ b,
                                  T *__restrict__ r, int i, int j, int k) {
    long id;
    const T v = resid_value(L, x, b, i, j, k, id);
    r[id] = v;
}

__device__ inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// restriction: adjoint of the prolongation / 2 per halved axis,
// 1-D weights (1/8, 3/8, 3/8, 1/8) with reflection at the walls
template <typename T>
__device__ inline T restrict_value(const LV<T> &F, int cx, int cy, int cz, const T *rf, int I, int J, int K) {
    const T W[4] = {(T)0.125, (T)0.375, (T)0.375, (T)0.125};
    T s = (T)0;
    int tk0 = cz ? -1 : 0, tk1 = cz ? 2 : 0;
    int tj0 = cy ? -1 : 0, tj1 = cy ? 2 : 0;
    int ti0 = cx ? -1 : 0, ti1 = cx ? 2 : 0;
    for (int tk = tk0; tk <= tk1; tk++) {
        int kf = cz ? clampi(2 * K + tk, 0, F.nz - 1) : K;
        T wk = cz ? W[tk + 1] : (T)1;
        for (int tj = tj0; tj <= tj1; tj++) {
            int jf = cy ? clampi(2 * J + tj, 0, F.ny - 1) : J;
            T wj = cy ? W[tj + 1] : (T)1;
            T sj = (T)0;
            for (int ti = ti0; ti <= ti1; ti++) {
                int ifn = cx ? clampi(2 * I + ti, 0, F.nx - 1) : I;
                T wi = cx ? W[ti + 1] : (T)1;
                sj += wi * rf[lidx(F, ifn, jf, kf)];
            }
            s += wk * wj * sj;
        }
    }
    return s;
}

template <typename T>
__device__ inline void restrict_cell(const LV<T> &F, const LV<T> &C, int cx, int cy, int cz, const T *__restrict__ rf,
                                     T *__restrict__ bc, int I, int J, int K) {
    const T s = restrict_value(F, cx, cy, cz, rf, I, J, K);
    bc[lidx(C, I, J, K)] = s;
}

// restriction with the halved axes fixed at compile time: the 4 x 4 x 4
// taps unroll, so their loads are issued together; same sums in the same
// order as restrict_cell, so the same bits
template <typename T, int CX, int CY, int CZ>
__device__ inline T restrict_sum(const LV<T> &F, const T *__restrict__ rf, int I, int J, int K) {
    const T W[4] = {(T)0.125, (T)0.375, (T)0.375, (T)0.125};
    T s = (T)0;
#pragma unroll
    for (int tk = (CZ ? -1 : 0); tk <= (CZ ? 2 : 0); tk++) {
        const int kf = CZ ? clampi(2 * K + tk, 0, F.nz - 1) : K;
        const T wk = CZ ? W[tk + 1] : (T)1;
#pragma unroll
        for (int tj = (CY ? -1 : 0); tj <= (CY ? 2 : 0); tj++) {
            const int jf = CY ? clampi(2 * J + tj, 0, F.ny - 1) : J;
            const T wj = CY ? W[tj + 1] : (T)1;
            T sj = (T)0;
#pragma unroll
            for (int ti = (CX ? -1 : 0); ti <= (CX ? 2 : 0); ti++) {
                const int ifn = CX ? clampi(2 * I + ti, 0, F.nx - 1) : I;
                const T wi = CX ? W[ti + 1] : (T)1;
                sj += wi * rf[lidx(F, ifn, jf, kf)];
            }
            s += wk * wj * sj;
        }
    }
    return s;
}

// prolongation: (3/4, 1/4) per halved axis, reflecting: the correction
// P x_c at fine cell (i, j, k), coarse values from xcv(I, J, K)
template <typename T, typename G>
__device__ inline T prolong_sum(const LV<T> &C, int cx, int cy, int cz, int i, int j, int k, G xcv) {
    int K0, K1, J0, J1, I0, I1;
    T wk1, wj1, wi1;
    if (cz) { K0 = k >> 1; K1 = clampi((k & 1) ? K0 + 1 : K0 - 1, 0, C.nz - 1); wk1 = (T)0.25; } else { K0 = K1 = k; wk1 = (T)0; }
    if (cy) { J0 = j >> 1; J1 = clampi((j & 1) ? J0 + 1 : J0 - 1, 0, C.ny - 1); wj1 = (T)0.25; } else { J0 = J1 = j; wj1 = (T)0; }
    if (cx) { I0 = i >> 1; I1 = clampi((i & 1) ? I0 + 1 : I0 - 1, 0, C.nx - 1); wi1 = (T)0.25; } else { I0 = I1 = i; wi1 = (T)0; }
    T wk0 = (T)1 - wk1, wj0 = (T)1 - wj1, wi0 = (T)1 - wi1;
    return wk0 * (wj0 * (wi0 * xcv(I0, J0, K0) + wi1 * xcv(I1, J0, K0)) +
                  wj1 * (wi0 * xcv(I0, J1, K0) + wi1 * xcv(I1, J1, K0))) +
           wk1 * (wj0 * (wi0 * xcv(I0, J0, K1) + wi1 * xcv(I1, J0, K1)) +
                  wj1 * (wi0 * xcv(I0, J1, K1) + wi1 * xcv(I1, J1, K1)));
}

// x_f = x_in + P x_c (x_in may be x_f itself)
template <typename T>
__device__ inline void prolong_cell(const LV<T> &F, const LV<T> &C, int cx, int cy, int cz, const T *__restrict__ xc,
                                    const T *xin, T *xf, int i, int j, int k) {
    const T s = prolong_sum(C, cx, cy, cz, i, j, k, [&](int I, int J, int K) { return xc[lidx(C, I, J, K)]; });
    const long id = lidx(F, i, j, k);
    xf[id] = xin[id] + s;
}

// coarsest level: x = Minv (b - mean b), then x -= mean x (whole block;
// blockDim.x >= n, n <= 1024); MinvT is the inverse stored transposed
template <typename T>
__device__ inline void coarse_solve(const T *__restrict__ Minv, int n, const T *__restrict__ b, T *__restrict__ x,
                                    double *sh, double *bb, double *mean_s) {
    double v[1] = {0.0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) v[0] += (double)b[a];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) *mean_s = v[0] / n;
    __syncthreads();
    double mb = *mean_s;
    for (int a = threadIdx.x; a < n; a += blockDim.x) bb[a] = (double)b[a] - mb;
    __syncthreads();
    double xr[1] = {0.0};
    double mine = 0.0;
    if ((int)threadIdx.x < n) {
        double s = 0.0;
        for (int cc = 0; cc < n; cc++) s += (double)Minv[(long)cc * n + threadIdx.x] * bb[cc];
        mine = s;
        xr[0] = s;
    }
    __syncthreads();
    block_sum<1>(xr, sh);
    if (threadIdx.x == 0) *mean_s = xr[0] / n;
    __syncthreads();
    if ((int)threadIdx.x < n) x[threadIdx.x] = (T)(mine - *mean_s);
    __syncthreads();
}

// red-black Gauss-Seidel half sweep; threads tile (ceil(nx/2), ny, nz)
template <typename T>
__global__ void __launch_bounds__(256) k_rbgs(LV<T> L, T *__restrict__ x, const T *__restrict__ b, int color,
                                              int zero_init, Tiles t) {
    int ih, j, k;
    if (!tile_point(t, ih, j, k)) return;
    int i = 2 * ih + ((j + k + color) & 1);
    if (i >= L.nx) return;
    rbgs_cell(L, x, b, i, j, k, zero_init != 0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_resid(LV<T> L, const T *__restrict__ x, const T *__restrict__ b,
                                                  T *__restrict__ r, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    resid_cell(L, x, b, r, i, j, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k_restrict(LV<T> F, LV<T> C, int cx, int cy, int cz,
                                                  const T *__restrict__ rf, T *__restrict__ bc, Tiles t) {
    int I, J, K;
    if (!tile_point(t, I, J, K)) return;
    restrict_cell(F, C, cx, cy, cz, rf, bc, I, J, K);
}

template <typename T, int CX, int CY, int CZ>
__global__ void __launch_bounds__(256) k_restrict_t(LV<T> F, LV<T> C, const T *__restrict__ rf, T *__restrict__ bc,
                                                    Tiles t) {
    int I, J, K;
    if (!tile_point(t, I, J, K)) return;
    bc[lidx(C, I, J, K)] = restrict_sum<T, CX, CY, CZ>(F, rf, I, J, K);
}

// Residual and restriction in one pass for the coarsening pattern of the
// large levels, y and z halved and x kept (cx, cy, cz) = (0, 1, 1): a thread
// owns coarse column (I, J) = fine column i = I, rows 2J-1 .. 2J+2, and walks
// its coarse planes upwards.  Each fine plane's residual is formed once, on
// the four rows, from x planes held in registers (rows 2J-2 .. 2J+3), summed
// over y with the restriction weights, and the z weights are applied over a
// window of four such plane sums; the residual array never reaches HBM.  The
// same weights and wall reflection as restrict_cell, summed y first.
constexpr int RR_KC = 8;   // coarse planes per block

template <typename T>
__global__ void __launch_bounds__(256) k_resid_restrict_011(LV<T> F, LV<T> C, const T *__restrict__ x,
                                                            const T *__restrict__ b, T *__restrict__ bc, KTiles t) {
    const int bid = blockIdx.x;
    const int lt = (bid % NXCD) * (int)t.per_xcd + bid / NXCD;
    if (lt >= t.n) return;
    const int I = (lt % t.tx) * TILE_X + threadIdx.x;
    const int rest = lt / t.tx;
    const int J = (rest % t.ty) * TILE_Y + threadIdx.y;
    const int K0 = (rest / t.ty) * RR_KC;
    const int K1 = K0 + RR_KC < C.nz ? K0 + RR_KC : C.nz;
    // every lane of a wave stays to the end: the x neighbours are shuffled
    const bool act = I < C.nx && J < C.ny;
    const int lane = threadIdx.x;   // blockDim.x == 64: one wave per row of the block
    const int i = act ? I : C.nx - 1;
    const int Jc = act ? J : C.ny - 1;
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const int pl = F.px * F.py;   // 32-bit offsets: the host checks the level size
    const T W0 = (T)0.125, W1 = (T)0.375;
    // register rows r6 = 0..5 hold fine rows 2J-2 .. 2J+3 (clamped)
    int roff[6];
#pragma unroll
    for (int r6 = 0; r6 < 6; r6++) roff[r6] = clampi(2 * Jc - 2 + r6, 0, ny - 1) * F.px + i;
    // planes kc-1 .. kc+2 of x and kc, kc+1 of b in registers: the loads of
    // a plane are issued two steps before its residual needs them.  The
    // wave's edge lanes also prefetch the x neighbour across the wave edge
    // (west for lane 0, east for lane 63) of rows 2J-1 .. 2J+2 of planes
    // kc .. kc+2, so no load waits inside the residual.  FP32 only: the
    // FP64 form runs faster with the registers left to occupancy and loads
    // those values where it uses them.
    constexpr bool PF = sizeof(T) == 4;
    T xm[6], xc[6], xp[6], xq[6], bc0[4], bc1[4], ec[4], ep[4], eq[4];
    const int eoff = !PF ? 0 : (lane == 0 && i > 0) ? -1 : ((lane == 63 && i < nx - 1) ? 1 : 0);
    int kc = 0;   // fine plane held in xc
    auto load_plane = [&](int kf, T *dst, T *edst) {
        if (kf < 0 || kf >= nz) return;
#pragma unroll
        for (int r6 = 0; r6 < 6; r6++) dst[r6] = x[kf * pl + roff[r6]];
        if (eoff != 0) {
#pragma unroll
            for (int tj = 0; tj < 4; tj++) edst[tj] = x[kf * pl + roff[1 + tj] + eoff];
        }
    };
    auto load_b = [&](int kf, T *dst) {
        if (kf < 0 || kf >= nz) return;
#pragma unroll
        for (int tj = 0; tj < 4; tj++) dst[tj] = b[kf * pl + roff[1 + tj]];
    };
    auto step = [&]() {
#pragma unroll
        for (int r6 = 0; r6 < 6; r6++) { xm[r6] = xc[r6]; xc[r6] = xp[r6]; xp[r6] = xq[r6]; }
#pragma unroll
        for (int tj = 0; tj < 4; tj++) { bc0[tj] = bc1[tj]; ec[tj] = ep[tj]; ep[tj] = eq[tj]; }
        kc++;
        load_plane(kc + 2, xq, eq);
        load_b(kc + 1, bc1);
    };
    // y-restricted residual of fine plane kc (absent planes are not read).
    // Register row r6 = 1 + tj holds fine row 2J-1+tj (clamped), so every
    // register index is static; the rows 2J-1 and 2J+2 that fall outside
    // the level reflect onto rows 2J and 2J+1, whose residuals they repeat.
    const bool lo_wall = 2 * Jc - 1 < 0, hi_wall = 2 * Jc + 2 > ny - 1;
    auto plane_sum = [&]() -> T {
        const bool hb = kc > 0, hf = kc < nz - 1;
        const bool hw = i > 0, he = i < nx - 1;
        T rs[4];
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            const int r6 = 1 + tj;
            const int jr = 2 * Jc - 1 + tj;   // unclamped row (the wall cases are replaced below)
            const bool hs = jr > 0, hn = jr < ny - 1;
            const T xcc = xc[r6];
            // x neighbours from the adjacent lanes (a wave is 64 consecutive
            // i); the wave's edge lanes take their prefetched values (ec[tj]
            // is row roff[1 + tj])
            T xw = __shfl_up(xcc, 1, 64), xe = __shfl_down(xcc, 1, 64);
            const int id = kc * pl + roff[r6];
            if (lane == 0) xw = hw ? (PF ? ec[tj] : x[id - 1]) : xcc;
            if (lane == 63) xe = he ? (PF ? ec[tj] : x[id + 1]) : xcc;
            xw = hw ? xw : xcc;
            xe = he ? xe : xcc;
            const T xs = hs ? xc[r6 - 1] : xcc, xn = hn ? xc[r6 + 1] : xcc;
            const T xb = hb ? xm[r6] : xcc, xf = hf ? xp[r6] : xcc;
            const T sr = F.cx * (xcc - xw) + F.cx * (xcc - xe) + F.cy * (xcc - xs) + F.cy * (xcc - xn) +
                         F.cz * (xcc - xb) + F.cz * (xcc - xf);
            rs[tj] = bc0[tj] - sr;   // bc0[tj]: b of the same row
        }
        if (lo_wall) rs[0] = rs[1];
        if (hi_wall) rs[3] = rs[2];
        T sum = (T)0;
#pragma unroll
        for (int tj = 0; tj < 4; tj++) sum += ((tj == 0 || tj == 3) ? W0 : W1) * rs[tj];
        return sum;
    };
    kc = 2 * K0 - 1 < 0 ? 0 : 2 * K0 - 1;
    T edummy[4];
    load_plane(kc - 1, xm, edummy);
    load_plane(kc, xc, ec);
    load_plane(kc + 1, xp, ep);
    load_plane(kc + 2, xq, eq);
    load_b(kc, bc0);
    load_b(kc + 1, bc1);
    T Rm1 = plane_sum();
    if (kc < 2 * K0) step();
    T R0 = plane_sum();
#pragma unroll 1
    for (int K = K0; K < K1; K++) {
        step();
        const T R1 = plane_sum();
        T R2 = R1;
        if (2 * K + 2 < nz) {
            step();
            R2 = plane_sum();
        }
        if (act) bc[lidx(C, I, J, K)] = W0 * Rm1 + W1 * R0 + W1 * R1 + W0 * R2;
        Rm1 = R1;
        R0 = R2;
    }
}

// Prolongation for the (0, 1, 1) pattern: a thread owns fine column (i, j)
// and walks PR_KC fine planes, keeping the coarse values of rows J0, J1 of
// the coarse planes it straddles in registers (one coarse load per two fine
// planes instead of four per cell).  Same expression as prolong_cell (its x
// weights are 1 and 0 here), so the result is bit-identical.
constexpr int PR_KC = 8;

template <typename T>
__global__ void __launch_bounds__(256) k_prolong_011(LV<T> F, LV<T> C, const T *__restrict__ xc, const T *xin,
                                                     T *xf, KTiles t) {
    int i, j, k0, k1;
    bool act;
    {
        const int bid = blockIdx.x;
        const int lt = (bid % NXCD) * (int)t.per_xcd + bid / NXCD;
        if (lt >= t.n) return;
        i = (lt % t.tx) * TILE_X + threadIdx.x;
        const int rest = lt / t.tx;
        j = (rest % t.ty) * TILE_Y + threadIdx.y;
        k0 = (rest / t.ty) * PR_KC;
        k1 = k0 + PR_KC < F.nz ? k0 + PR_KC : F.nz;
        act = i < F.nx && j < F.ny;
    }
    if (!act) return;
    const int J0 = j >> 1, J1 = clampi((j & 1) ? J0 + 1 : J0 - 1, 0, C.ny - 1);
    const T wj1 = (T)0.25, wj0 = (T)1 - wj1, wk1 = (T)0.25, wk0 = (T)1 - wk1;
    const int cpl = C.px * C.py, fpl = F.px * F.py;
    const int c0 = J0 * C.px + i, c1 = J1 * C.px + i;
    // coarse planes Kc-1, Kc, Kc+1 (clamped) of rows J0 (a) and J1 (b)
    int Kc = k0 >> 1;
    T am, a0, ap, bm, b0, bp;
    {
        const int Km = Kc > 0 ? Kc - 1 : 0, Kp = Kc + 1 < C.nz ? Kc + 1 : C.nz - 1;
        am = xc[Km * cpl + c0]; a0 = xc[Kc * cpl + c0]; ap = xc[Kp * cpl + c0];
        bm = xc[Km * cpl + c1]; b0 = xc[Kc * cpl + c1]; bp = xc[Kp * cpl + c1];
    }
    for (int k = k0; k < k1; k++) {
        if ((k >> 1) != Kc) {
            Kc++;
            am = a0; a0 = ap; bm = b0; b0 = bp;
            const int Kp = Kc + 1 < C.nz ? Kc + 1 : C.nz - 1;
            ap = xc[Kp * cpl + c0];
            bp = xc[Kp * cpl + c1];
        }
        const T a1 = (k & 1) ? ap : am, b1 = (k & 1) ? bp : bm;
        const T s = wk0 * (wj0 * a0 + wj1 * b0) + wk1 * (wj0 * a1 + wj1 * b1);
        const int id = k * fpl + j * F.px + i;
        xf[id] = xin[id] + s;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prolong(LV<T> F, LV<T> C, int cx, int cy, int cz, const T *__restrict__ xc,
                                                 const T *xin, T *xf, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    prolong_cell(F, C, cx, cy, cz, xc, xin, xf, i, j, k);
}

// One red-black sweep pair on a large level in a single pass, xn =
// second(first(xo)), first colour = (i + j + k) % 2 == F (the pre-smoother
// sweeps colour 0 first, the post-smoother colour 1).  The first colour reads
// only second-colour neighbours of xo and the second only first-colour
// values of this pair, so the result equals two half sweeps bit for bit
// (same expression, rb_value) while xo and b cross HBM once.  ZERO: xo is
// the zero initial guess and is not read.  The block layout is the one of
// k_helm_gs_pair: a 64 x 16 column tile walking PAIR_TK planes through LDS
// rings of three planes of xo (halo 2) and of first-colour values (halo 1),
// xo and b fetched into registers two planes ahead (every block of the
// finest level is resident at once, so the plane-to-plane chain is the
// latency to hide), two barriers per plane (plane m-2 is stored after the
// first).
constexpr int PAIR_TX = 64, PAIR_TY = 16;
constexpr int PAIR_OX = PAIR_TX + 4, PAIR_OY = PAIR_TY + 4;
constexpr int PAIR_HX = PAIR_TX + 2, PAIR_HY = PAIR_TY + 2;
constexpr int PAIR_THREADS = 512;
constexpr int PAIR_ORING = 3;   // the second-colour pass reads no xo, so staging may overwrite plane m-1
constexpr int PAIR_OLD_PER = (PAIR_OY * PAIR_OX + PAIR_THREADS - 1) / PAIR_THREADS;
constexpr int PAIR_RED = PAIR_HY * (PAIR_HX / 2);
constexpr int PAIR_RED_PER = (PAIR_RED + PAIR_THREADS - 1) / PAIR_THREADS;
static_assert(PAIR_TX * PAIR_TY == 2 * PAIR_THREADS, "one second-colour cell per thread");
// Levels at least this large use the fused pair (smaller ones have too
// few tiles to hide the plane-by-plane chain of a block); a block walks 16
// planes on levels of 8 M cells or more, 8 below (twice the blocks)
constexpr long PAIR_LEVEL_CELLS = 4L << 20;

static int pair_chunk(const Level &L) { return (long)L.nx * L.ny * L.nz >= (8L << 20) ? 16 : 8; }

struct PairTiles {
    int tx, ty, kc, tk;   // tk: planes per block (the chain a block walks)
    int n, per_xcd;
};

static PairTiles make_pair_tiles(int nx, int ny, int nz, int tk) {
    PairTiles t;
    t.tx = (nx + PAIR_TX - 1) / PAIR_TX;
    t.ty = (ny + PAIR_TY - 1) / PAIR_TY;
    t.tk = tk;
    t.kc = (nz + tk - 1) / tk;
    t.n = t.tx * t.ty * t.kc;
    t.per_xcd = (t.n + NXCD - 1) / NXCD;
    return t;
}

// DOT: the block's partial sum of b . xn over its own cells (the final
// values of both colours) goes to part[blockIdx.x] -- r . z of a V-cycle
// whose last sweep pair this is.  PROL (levels coarsened by (0, 1, 1)): the
// pair starts from xo + P xc, the coarse correction prolonged as
// k_prolong_011 does it (same expression, so the same bits), formed as each
// plane is staged instead of in a pass of its own.
template <typename T, int F, bool ZERO, bool DOT, bool PROL = false>
__global__ void __launch_bounds__(PAIR_THREADS, sizeof(T) == 4 ? ((PROL && !DOT) ? 8 : 4) : 2)
    k_rbgs_pair(LV<T> L, const T *__restrict__ xo, T *__restrict__ xn, const T *__restrict__ b, PairTiles P,
                double *__restrict__ part, LV<T> C = LV<T>(), const T *__restrict__ xc = nullptr) {
    __shared__ T so[PAIR_ORING][PAIR_OY * PAIR_OX];
    __shared__ T sr[3][PAIR_HY * PAIR_HX];
    const int nx = L.nx, ny = L.ny, nz = L.nz, px = L.px;
    const int plane = L.px * L.py;
    const int bid = blockIdx.x;
    const int lt = (bid % NXCD) * P.per_xcd + bid / NXCD;
    if (lt >= P.n) {                                        // whole block
        if (DOT && threadIdx.x == 0) part[bid] = 0.0;
        return;
    }
    double dacc[1] = {0.0};
    const int tx = lt % P.tx, rest = lt / P.tx;
    const int i0 = tx * PAIR_TX, j0 = (rest % P.ty) * PAIR_TY;
    const int k0 = (rest / P.ty) * P.tk;
    const int k1 = k0 + P.tk < nz ? k0 + P.tk : nz;
    const int t = threadIdx.x;

    for (int e = t; e < PAIR_ORING * PAIR_OY * PAIR_OX; e += PAIR_THREADS) (&so[0][0])[e] = (T)0;
    for (int e = t; e < 3 * PAIR_HY * PAIR_HX; e += PAIR_THREADS) (&sr[0][0])[e] = (T)0;
    __syncthreads();

    int ooff[PAIR_OLD_PER];
#pragma unroll
    for (int u = 0; u < PAIR_OLD_PER; u++) {
        const int e = t + u * PAIR_THREADS;
        const int r = e / PAIR_OX, cc = e - r * PAIR_OX;
        const int i = i0 - 2 + cc, j = j0 - 2 + r;
        ooff[u] = (e < PAIR_OY * PAIR_OX && i >= 0 && i < nx && j >= 0 && j < ny) ? j * px + i : -1;
    }
    T ro[PAIR_OLD_PER], ro2[PAIR_OLD_PER];   // xo of the next two planes to stage
    // PROL: the coarse values of each staged point, rows J0 (a) and J1 (b)
    // of coarse planes K0 = m / 2 (0) and its neighbour K1 (1), fetched one
    // plane ahead of their use (the xo values two)
    T ca0[PAIR_OLD_PER], cb0[PAIR_OLD_PER], ca1[PAIR_OLD_PER], cb1[PAIR_OLD_PER];
    int coff0[PAIR_OLD_PER], coff1[PAIR_OLD_PER];
    const int cpl = C.px * C.py;
#pragma unroll
    for (int u = 0; u < PAIR_OLD_PER; u++) {
        ro2[u] = (T)0;
        ca0[u] = cb0[u] = ca1[u] = cb1[u] = (T)0;
        coff0[u] = coff1[u] = 0;
        if (PROL && ooff[u] >= 0) {
            const int e = t + u * PAIR_THREADS;
            const int r = e / PAIR_OX, cc = e - r * PAIR_OX;
            const int i = i0 - 2 + cc, j = j0 - 2 + r;
            const int J0 = j >> 1, J1 = clampi((j & 1) ? J0 + 1 : J0 - 1, 0, C.ny - 1);
            coff0[u] = J0 * C.px + i;
            coff1[u] = J1 * C.px + i;
        }
    }
    // staging on named register arrays (the compiler keeps them in
    // registers only while no pointer to them is taken): fetch the xo plane
    // after next into ro2, shift it to ro, fetch its coarse values, store
    // ro's plane into its LDS slot
    auto fetch_next = [&](int m) {
#pragma unroll
        for (int u = 0; u < PAIR_OLD_PER; u++) ro2[u] = ooff[u] >= 0 ? xo[m * plane + ooff[u]] : (T)0;
    };
    auto shift = [&]() {
#pragma unroll
        for (int u = 0; u < PAIR_OLD_PER; u++) ro[u] = ro2[u];
    };
    auto fetch_coarse = [&](int m) {
        if (!PROL) return;
        const int K0 = m >> 1;
        const int K1 = (m & 1) ? (K0 + 1 < C.nz ? K0 + 1 : C.nz - 1) : (K0 > 0 ? K0 - 1 : 0);
#pragma unroll
        for (int u = 0; u < PAIR_OLD_PER; u++) {   // offset 0 outside the box: no branch around the loads
            ca0[u] = xc[K0 * cpl + coff0[u]];
            cb0[u] = xc[K0 * cpl + coff1[u]];
            ca1[u] = xc[K1 * cpl + coff0[u]];
            cb1[u] = xc[K1 * cpl + coff1[u]];
        }
    };
    auto put_cur = [&](int m) {
        T *S = so[m % PAIR_ORING];
        const T wj1 = (T)0.25, wj0 = (T)1 - wj1, wk1 = (T)0.25, wk0 = (T)1 - wk1;
#pragma unroll
        for (int u = 0; u < PAIR_OLD_PER; u++) {
            const int e = t + u * PAIR_THREADS;
            if (e < PAIR_OY * PAIR_OX) {
                T v = ro[u];
                if (PROL) {
                    const T s = wk0 * (wj0 * ca0[u] + wj1 * cb0[u]) + wk1 * (wj0 * ca1[u] + wj1 * cb1[u]);
                    v = ooff[u] >= 0 ? v + s : (T)0;
                }
                S[e] = v;
            }
        }
    };
    int rrow[PAIR_RED_PER], rc2[PAIR_RED_PER];
#pragma unroll
    for (int u = 0; u < PAIR_RED_PER; u++) {
        const int e = t + u * PAIR_THREADS;
        rrow[u] = e < PAIR_RED ? e / (PAIR_HX / 2) : 0;
        rc2[u] = e < PAIR_RED ? e - rrow[u] * (PAIR_HX / 2) : -1;
    }
    // b of this thread's first-colour cells of plane m (valid: inside the box)
    T rb[PAIR_RED_PER];
    bool rv[PAIR_RED_PER];
    auto fetch_red = [&](int m, T *b_, bool *v_) {
#pragma unroll
        for (int u = 0; u < PAIR_RED_PER; u++) {
            const int j = j0 - 1 + rrow[u];
            const int i = i0 - 1 + 2 * rc2[u] + ((F + 1 + j + m) & 1);
            v_[u] = rc2[u] >= 0 && i >= 0 && i < nx && j >= 0 && j < ny;
            b_[u] = v_[u] ? b[m * plane + j * px + i] : (T)0;
        }
    };
#pragma unroll
    for (int u = 0; u < PAIR_RED_PER; u++) { rb[u] = (T)0; rv[u] = false; }
    const int br = t / (PAIR_TX / 2), bc2 = t % (PAIR_TX / 2);
    const int bj = j0 + br;
    T bb = (T)0;
    auto fetch_black = [&](int k) -> T {
        const int i = i0 + 2 * bc2 + ((1 + F + bj + k) & 1);
        return (i < nx && bj < ny) ? b[k * plane + bj * px + i] : (T)0;
    };

    // both colours of plane k, from its sr slot to xn (coalesced rows)
    auto write_plane = [&](int k) {
        if (k < k0 || k >= k1) return;
        const T *R = sr[k % 3];
        for (int e = t; e < PAIR_TY * PAIR_TX; e += PAIR_THREADS) {
            const int r = e / PAIR_TX, c2 = e - r * PAIR_TX;
            const int ii = i0 + c2, jj = j0 + r;
            if (ii < nx && jj < ny) xn[k * plane + jj * px + ii] = R[(r + 1) * PAIR_HX + c2 + 1];
        }
    };

    // b of the first-colour cells of planes m + 1 and m + 2 (rb: plane m),
    // and the second-colour b of planes k + 1 and k + 2 (bb: k = m - 1)
    T rb_n[PAIR_RED_PER];
    bool rv_n[PAIR_RED_PER];
#pragma unroll
    for (int u = 0; u < PAIR_RED_PER; u++) { rb_n[u] = (T)0; rv_n[u] = false; }
    T bb_n = (T)0;
    if (!ZERO) {
        if (k0 >= 2) { fetch_next(k0 - 2); shift(); fetch_coarse(k0 - 2); put_cur(k0 - 2); }
        if (k0 >= 1) { fetch_next(k0 - 1); shift(); fetch_coarse(k0 - 1); put_cur(k0 - 1); }
        fetch_next(k0);
        shift();
        fetch_coarse(k0);
        if (k0 + 1 < nz) fetch_next(k0 + 1);
    }
    if (k0 >= 1) fetch_red(k0 - 1, rb, rv);
    if (k0 < nz) fetch_red(k0, rb_n, rv_n);
    for (int m = k0 - 1; m <= k1; m++) {
        if (!ZERO) {
            if (m + 1 < nz) put_cur(m + 1);
            shift();
            if (m + 2 < nz) fetch_coarse(m + 2);
            if (m + 3 < nz) fetch_next(m + 3);
        }
        T rb_nn[PAIR_RED_PER];
        bool rv_nn[PAIR_RED_PER];
#pragma unroll
        for (int u = 0; u < PAIR_RED_PER; u++) { rb_nn[u] = (T)0; rv_nn[u] = false; }
        if (m + 2 < nz) fetch_red(m + 2, rb_nn, rv_nn);
        const T bb_nn = (m + 1 >= k0 && m + 1 < k1) ? fetch_black(m + 1) : (T)0;
        __syncthreads();
        write_plane(m - 2);   // final; its sr slot is next written by the first-colour pass of m + 1
        if (m >= 0 && m < nz) {
            const T *S = so[m % PAIR_ORING], *Sb = so[(m + PAIR_ORING - 1) % PAIR_ORING],
                    *Sf = so[(m + 1) % PAIR_ORING];
            T *R = sr[m % 3];
#pragma unroll
            for (int u = 0; u < PAIR_RED_PER; u++) {
                if (!rv[u]) continue;
                const int r = rrow[u];
                const int j = j0 - 1 + r;
                const int cc = 2 * rc2[u] + ((F + 1 + j + m) & 1);
                const int i = i0 - 1 + cc;
                const int o = (r + 1) * PAIR_OX + cc + 1;
                const T v = rb_value(L, i, j, m, rb[u], S[o - 1], S[o + 1], S[o - PAIR_OX], S[o + PAIR_OX], Sb[o],
                                     Sf[o]);
                R[r * PAIR_HX + cc] = v;
                if (DOT && m >= k0 && m < k1 && i >= i0 && i < i0 + PAIR_TX && j >= j0 && j < j0 + PAIR_TY)
                    dacc[0] += (double)v * (double)rb[u];
            }
        }
        __syncthreads();
        const int k = m - 1;
        const int cc = 2 * bc2 + ((1 + F + bj + k) & 1);
        const int i = i0 + cc;
        if (k >= k0 && k < k1 && i < nx && bj < ny) {
            T *R = sr[k % 3];
            const T *Rb = sr[(k + 2) % 3], *Rf = sr[(k + 1) % 3];
            const int h = (br + 1) * PAIR_HX + cc + 1;
            const T v = rb_value(L, i, bj, k, bb, R[h - 1], R[h + 1], R[h - PAIR_HX], R[h + PAIR_HX], Rb[h], Rf[h]);
            R[h] = v;
            if (DOT) dacc[0] += (double)v * (double)bb;
        }
        // no barrier: the next staging writes the xo slot of plane m - 1,
        // which the second-colour pass does not read
#pragma unroll
        for (int u = 0; u < PAIR_RED_PER; u++) {
            rb[u] = rb_n[u]; rv[u] = rv_n[u];
            rb_n[u] = rb_nn[u]; rv_n[u] = rv_nn[u];
        }
        bb = bb_n;
        bb_n = bb_nn;
    }
    __syncthreads();
    write_plane(k1 - 1);
    if (DOT) {
        __shared__ double dsh[16];
        block_sum<1>(dacc, dsh);
        if (threadIdx.x == 0) part[bid] = dacc[0];
    }
}

template <typename T>
__global__ void k_coarse(const T *__restrict__ Minv, int n, const T *__restrict__ b, T *__restrict__ x) {
    __shared__ double sh[16];
    __shared__ double bb[1024];
    __shared__ double mean_s;
    coarse_solve(Minv, n, b, x, sh, bb, &mean_s);
}

// The small levels of a V-cycle run in one workgroup of 1024 threads with
// every level resident in LDS (compact layouts), one barrier between phases:
// their launches (~3-5 us each, 11 per level) and L2 round trips would
// otherwise cost far more than their work.  Level l0 is read from (b) and
// written to (x) global memory once; one residual buffer serves all levels.
// The first level run this way is the largest one that has at most
// SMALL_LEVEL_CELLS cells and whose levels fit SMALL_LDS_BYTES.  One CU is
// the limit: an 8192-cell level costs ~40 us there against ~22 us as blocked
// launches over the chip (measured at 512x256x128), so the cycle starts at
// 1024 cells.
constexpr long SMALL_LEVEL_CELLS = 1024;
constexpr int SMALL_LDS_BYTES = 156 * 1024;
constexpr int MAXLEV = 16;

template <typename T>
struct SmallCycle {
    int l0, nl, pre, post, coarse_n;
    LV<T> lv[MAXLEV];          // compact layouts of the LDS copies
    LV<T> g0;                  // layout of level l0 in global memory
    T *gx0;                    // level l0 in global memory: x out
    const T *gb0;              //                            b in
    int xo[MAXLEV], bo[MAXLEV], ro;   // LDS offsets in elements of T
    int cx[MAXLEV], cy[MAXLEV], cz[MAXLEV];
    const T *Minv;
};

// LDS bytes of a small cycle starting at level l0: coarse-solve scratch
// (1024 + 32 doubles), then x and b of every level and one residual buffer
template <typename T>
static int small_lds_bytes(const tb_ctx *c, int l0) {
    long e = 0;
    for (int m = l0; m < c->nlev; m++) e += 2L * c->lev[m].nx * c->lev[m].ny * c->lev[m].nz;
    e += (long)c->lev[l0].nx * c->lev[l0].ny * c->lev[l0].nz;
    return (int)(sizeof(double) * (1024 + 32) + sizeof(T) * e);
}

// a / d and a % d for 0 <= a < 2^22, d >= 1, from the float reciprocal inv
// of d: the float quotient is within 1 of the true one, and one correction
// step fixes it (a handful of instructions where an integer division takes
// dozens, which every thread of a 1024-thread block would pay)
__device__ inline int divmod_small(int a, int d, float inv, int &rem) {
    int q = (int)((float)a * inv);
    int r = a - q * d;
    if (r < 0) { q--; r += d; }
    else if (r >= d) { q++; r -= d; }
    rem = r;
    return q;
}

// Calls f(i, j, k) for this thread's cells of an nx x ny x nz box, cell q =
// threadIdx.x + s * blockDim.x in x-fastest order; the position advances by
// carries instead of a division per cell.  Boxes have fewer than 2^22 cells.
template <typename F>
__device__ inline void block_cells(int nx, int ny, int nz, F f) {
    const int n = nx * ny * nz;
    const float inx = 1.0f / (float)nx, iny = 1.0f / (float)ny;
    int q = threadIdx.x;
    int i, j;
    const int r = divmod_small(q, nx, inx, i);
    int k = divmod_small(r, ny, iny, j);
    int di, dj;
    const int dr = divmod_small((int)blockDim.x, nx, inx, di);
    const int dk = divmod_small(dr, ny, iny, dj);
    for (; q < n; q += blockDim.x) {
        f(i, j, k);
        i += di;
        j += dj;
        k += dk;
        if (i >= nx) { i -= nx; j++; }
        if (j >= ny) { j -= ny; k++; }
    }
}

template <typename T>
__device__ inline void block_sweep(const LV<T> &L, T *x, const T *b, int color, bool zero_init) {
    block_cells((L.nx + 1) / 2, L.ny, L.nz, [&](int ih, int j, int k) {
        const int i = 2 * ih + ((j + k + color) & 1);
        if (i < L.nx) rbgs_cell(L, x, b, i, j, k, zero_init);
    });
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(1024) k_small_cycle(SmallCycle<T> A) {
    extern __shared__ double smem[];
    double *bb = smem, *sh = smem + 1024, *mean_s = smem + 1024 + 16;
    T *S = reinterpret_cast<T *>(smem + 1024 + 32);
    const int last = A.nl - 1;
    {
        const LV<T> &L = A.lv[A.l0];
        T *b0 = S + A.bo[A.l0];
        block_cells(L.nx, L.ny, L.nz, [&](int i, int j, int k) {
            b0[(k * L.ny + j) * L.nx + i] = A.gb0[lidx(A.g0, i, j, k)];
        });
        __syncthreads();
    }
    T *r = S + A.ro;
    for (int l = A.l0; l < last; l++) {
        const LV<T> &L = A.lv[l];
        T *x = S + A.xo[l];
        const T *b = S + A.bo[l];
        const int n = L.nx * L.ny * L.nz;
        if (A.pre > 0) {
            block_sweep(L, x, b, 0, true);
            block_sweep(L, x, b, 1, false);
            for (int s = 1; s < A.pre; s++) {
                block_sweep(L, x, b, 0, false);
                block_sweep(L, x, b, 1, false);
            }
        } else {
            for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] = (T)0;
            __syncthreads();
        }
        block_cells(L.nx, L.ny, L.nz, [&](int i, int j, int k) { resid_cell(L, (const T *)x, b, r, i, j, k); });
        __syncthreads();
        const LV<T> &C = A.lv[l + 1];
        T *bc = S + A.bo[l + 1];
        block_cells(C.nx, C.ny, C.nz, [&](int I, int J, int K) {
            restrict_cell(L, C, A.cx[l], A.cy[l], A.cz[l], (const T *)r, bc, I, J, K);
        });
        __syncthreads();
    }
    coarse_solve(A.Minv, A.coarse_n, (const T *)(S + A.bo[last]), S + A.xo[last], sh, bb, mean_s);
    for (int l = last - 1; l >= A.l0; l--) {
        const LV<T> &L = A.lv[l];
        const LV<T> &C = A.lv[l + 1];
        T *x = S + A.xo[l];
        const T *b = S + A.bo[l];
        block_cells(L.nx, L.ny, L.nz, [&](int i, int j, int k) {
            prolong_cell(L, C, A.cx[l], A.cy[l], A.cz[l], (const T *)(S + A.xo[l + 1]), (const T *)x, x, i, j, k);
        });
        __syncthreads();
        for (int s = 0; s < A.post; s++) {
            block_sweep(L, x, b, 1, false);
            block_sweep(L, x, b, 0, false);
        }
    }
    {
        const LV<T> &L = A.lv[A.l0];
        const T *x0 = S + A.xo[A.l0];
        block_cells(L.nx, L.ny, L.nz, [&](int i, int j, int k) {
            A.gx0[lidx(A.g0, i, j, k)] = x0[(k * L.ny + j) * L.nx + i];
        });
    }
}

// Mid-level smoothing in one launch.  Each block owns an output box of the
// level and runs NP red-black sweep pairs on it in LDS, recomputing the
// 2*NP-cell halo its last half sweep depends on, so no block reads what
// another writes: src and dst are distinct buffers.  Half sweep m = 1 ..
// 2*NP has colour FIRST for odd m and 1 - FIRST for even m and covers the
// box grown by 2*NP - m cells, which is every cell a later half sweep of the
// box reads.  ZERO: the sweeps start from x = 0 and read only b (the
// pre-smoother).  PROL: the starting iterate is src + P xc, the coarse
// level's correction prolonged on the fly (the post-smoother).  Every cell
// value is formed by the expressions of rb_value and prolong_sum, so the
// result equals the launch-per-half-sweep sequence bit for bit.
constexpr int BLK_THREADS = 1024;
constexpr int BLK_LDS_PAIR = 80 * 1024;      // two blocks per CU
constexpr int BLK_LDS_BYTES = 150 * 1024;    // one block per CU
constexpr int BLK_MIN_BOXES = 128;
// Above this many cells the halo recomputation (about 5x the cells of a box
// for two pairs) makes the blocked launch slower than its half sweeps,
// whose ~5 us launches it exists to save (measured at 512x256x128: the
// 1M-cell level 62 us blocked against 47, the 65k-cell level 17 against 42).
constexpr long BLK_MAX_CELLS = 1L << 18;

struct BlkShape {
    int bx, by, bz;          // output box
    int nbx, nby, nb;        // boxes along x, along y, in total
    int per_xcd;             // ceil(nb / 8)
    int cnx, cny, cnz;       // extents of the LDS copy of the coarse region (PROL)
};

template <typename T, int NP, bool ZERO, bool PROL, int FIRST>
__global__ void __launch_bounds__(BLK_THREADS) k_rbgs_blk(LV<T> L, const T *__restrict__ src, const T *__restrict__ b,
                                                          T *__restrict__ dst, BlkShape S, LV<T> C, int cx, int cy,
                                                          int cz, const T *__restrict__ xc) {
    extern __shared__ double smem_blk[];
    constexpr int H = 2 * NP;
    const int rx = S.bx + 2 * H, ry = S.by + 2 * H, rz = S.bz + 2 * H;
    const int rn = rx * ry * rz, rpl = rx * ry;
    T *xs = reinterpret_cast<T *>(smem_blk);
    T *bs = xs + rn;
    T *cs = bs + rn;
    // consecutive boxes (which share halos) on one XCD
    const int lt = (int)(blockIdx.x % NXCD) * S.per_xcd + (int)(blockIdx.x / NXCD);
    if (lt >= S.nb) return;
    const int ox = (lt % S.nbx) * S.bx, oy = ((lt / S.nbx) % S.nby) * S.by, oz = (lt / (S.nbx * S.nby)) * S.bz;
    const int gx = ox - H, gy = oy - H, gz = oz - H;     // level index of LDS cell (0, 0, 0)
    const int x0 = max(gx, 0), x1 = min(gx + rx, L.nx);
    const int y0 = max(gy, 0), y1 = min(gy + ry, L.ny);
    const int z0 = max(gz, 0), z1 = min(gz + rz, L.nz);
    int ci = 0, cj = 0, ck = 0;
    if (PROL) {
        // the coarse cells the region's prolongation reads
        ci = cx ? max((x0 >> 1) - 1, 0) : x0;
        cj = cy ? max((y0 >> 1) - 1, 0) : y0;
        ck = cz ? max((z0 >> 1) - 1, 0) : z0;
        const int ci1 = cx ? min(((x1 - 1) >> 1) + 1, C.nx - 1) : x1 - 1;
        const int cj1 = cy ? min(((y1 - 1) >> 1) + 1, C.ny - 1) : y1 - 1;
        const int ck1 = cz ? min(((z1 - 1) >> 1) + 1, C.nz - 1) : z1 - 1;
        block_cells(ci1 - ci + 1, cj1 - cj + 1, ck1 - ck + 1, [&](int a, int e, int f) {
            cs[(f * S.cny + e) * S.cnx + a] = xc[lidx(C, ci + a, cj + e, ck + f)];
        });
        __syncthreads();
    }
    block_cells(x1 - x0, y1 - y0, z1 - z0, [&](int a, int e, int f) {
        const int i = x0 + a, j = y0 + e, k = z0 + f;
        const int q = ((k - gz) * ry + (j - gy)) * rx + (i - gx);
        const long g = lidx(L, i, j, k);
        bs[q] = b[g];
        if (!ZERO) {
            T v = src[g];
            if (PROL) {
                const T p = prolong_sum(C, cx, cy, cz, i, j, k, [&](int I, int J, int K) {
                    return cs[((K - ck) * S.cny + (J - cj)) * S.cnx + (I - ci)];
                });
                v = v + p;
            }
            xs[q] = v;
        }
    });
    __syncthreads();
#pragma unroll
    for (int m = 1; m <= 2 * NP; m++) {
        const int color = (m & 1) ? FIRST : 1 - FIRST;
        const bool zero_init = ZERO && m == 1;
        const int e = 2 * NP - m;
        const int ax0 = max(ox - e, 0), ax1 = min(ox + S.bx + e, L.nx);
        const int ay0 = max(oy - e, 0), ay1 = min(oy + S.by + e, L.ny);
        const int az0 = max(oz - e, 0), az1 = min(oz + S.bz + e, L.nz);
        block_cells((ax1 - ax0 + 1) / 2, ay1 - ay0, az1 - az0, [&](int ih, int e2, int f) {
            const int j = ay0 + e2, k = az0 + f;
            const int i = ax0 + 2 * ih + ((ax0 + j + k + color) & 1);
            if (i >= ax1) return;
            const int q = ((k - gz) * ry + (j - gy)) * rx + (i - gx);
            const bool hw = i > 0, he = i < L.nx - 1, hs = j > 0, hn = j < L.ny - 1, hb = k > 0, hf = k < L.nz - 1;
            T xw = (T)0, xe = (T)0, xsv = (T)0, xn = (T)0, xb = (T)0, xf = (T)0;
            if (!zero_init) {
                xw = xs[q - (hw ? 1 : 0)]; xe = xs[q + (he ? 1 : 0)];
                xsv = xs[q - (hs ? rx : 0)]; xn = xs[q + (hn ? rx : 0)];
                xb = xs[q - (hb ? rpl : 0)]; xf = xs[q + (hf ? rpl : 0)];
            }
            xs[q] = rb_value(L, i, j, k, bs[q], xw, xe, xsv, xn, xb, xf);
        });
        __syncthreads();
    }
    block_cells(min(S.bx, L.nx - ox), min(S.by, L.ny - oy), min(S.bz, L.nz - oz), [&](int a, int e, int f) {
        const int i = ox + a, j = oy + e, k = oz + f;
        dst[lidx(L, i, j, k)] = xs[((k - gz) * ry + (j - gy)) * rx + (i - gx)];
    });
}

// LDS bytes of k_rbgs_blk for box (bx, by, bz) on level L (coarsened by its
// cx, cy, cz): x and b over the box grown by 2*np, and the coarse region of
// the prolongation (a grown extent r covers at most (r + 1) / 2 + 3 coarse
// cells along a halved axis)
static long blk_lds(const Level &L, int np, size_t tsize, const int *bx, int *cn) {
    const int h = 2 * np;
    const int r[3] = {bx[0] + 2 * h, bx[1] + 2 * h, bx[2] + 2 * h};
    const int c[3] = {L.cx, L.cy, L.cz};
    for (int a = 0; a < 3; a++) cn[a] = c[a] ? (r[a] + 1) / 2 + 3 : r[a];
    return (long)tsize * (2L * r[0] * r[1] * r[2] + (long)cn[0] * cn[1] * cn[2]);
}

// box of the mid-level smoother: start from 32 x 16 x 8 and halve the
// largest side (not below 8) until the LDS fits two blocks per CU (else one)
// and there are at least BLK_MIN_BOXES boxes; false if no box fits
static bool blk_shape_cap(const Level &L, int np, size_t tsize, long cap, BlkShape &S, int &lds);
static bool blk_shape(const Level &L, int np, size_t tsize, BlkShape &S, int &lds) {
    return blk_shape_cap(L, np, tsize, BLK_LDS_PAIR, S, lds) || blk_shape_cap(L, np, tsize, BLK_LDS_BYTES, S, lds);
}
static bool blk_shape_cap(const Level &L, int np, size_t tsize, long cap, BlkShape &S, int &lds) {
    int bx[3] = {32, 16, 8}, cn[3];
    const int n[3] = {L.nx, L.ny, L.nz};
    auto boxes = [&]() {
        return (long)((n[0] + bx[0] - 1) / bx[0]) * ((n[1] + bx[1] - 1) / bx[1]) * ((n[2] + bx[2] - 1) / bx[2]);
    };
    for (;;) {
        const bool big = blk_lds(L, np, tsize, bx, cn) > cap;
        if (!big && boxes() >= BLK_MIN_BOXES) break;
        int a = 0;   // the largest side, the later axis on a tie (long x rows)
        for (int q = 1; q < 3; q++)
            if (bx[q] >= bx[a]) a = q;
        if (bx[a] <= 8) {
            if (big) return false;
            break;
        }
        bx[a] /= 2;
    }
    const long bytes = blk_lds(L, np, tsize, bx, cn);
    S.bx = bx[0]; S.by = bx[1]; S.bz = bx[2];
    S.nbx = (n[0] + bx[0] - 1) / bx[0];
    S.nby = (n[1] + bx[1] - 1) / bx[1];
    S.nb = (int)boxes();
    S.per_xcd = (S.nb + NXCD - 1) / NXCD;
    S.cnx = cn[0]; S.cny = cn[1]; S.cnz = cn[2];
    lds = (int)bytes;
    return true;
}

template <typename T>
static LV<T> lv_of(const Level &L) {
    LV<T> v;
    v.nx = L.nx; v.ny = L.ny; v.nz = L.nz;
    v.px = L.px; v.py = (int)(L.plane / L.px);
    v.cx = (T)(1.0 / (L.hx * L.hx));
    v.cy = (T)(1.0 / (L.hy * L.hy));
    v.cz = (T)(1.0 / (L.hz * L.hz));
    return v;
}

static int choose(const Level &L, int &cx, int &cy, int &cz) {
    int n[3] = {L.nx, L.ny, L.nz};
    double h[3] = {L.hx, L.hy, L.hz};
    double hmax = h[0];
    if (h[1] > hmax) hmax = h[1];
    if (h[2] > hmax) hmax = h[2];
    int cc[3] = {0, 0, 0}, any = 0;
    for (int a = 0; a < 3; a++) {
        int can = (n[a] % 2 == 0) && (n[a] >= 4);
        if (can && 2.0 * h[a] <= hmax * (1.0 + 1e-9)) { cc[a] = 1; any = 1; }
    }
    // spacings within a factor 2 of each other: full coarsening
    if (!any)
        for (int a = 0; a < 3; a++)
            if ((n[a] % 2 == 0) && (n[a] >= 4)) { cc[a] = 1; any = 1; }
    cx = cc[0]; cy = cc[1]; cz = cc[2];
    return any;
}

static void level_apply_host(const Level &L, const double *x, double *y) {
    double cx = 1.0 / (L.hx * L.hx), cy = 1.0 / (L.hy * L.hy), cz = 1.0 / (L.hz * L.hz);
    long pl = (long)L.nx * L.ny;
    for (int k = 0; k < L.nz; k++)
        for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++) {
                long id = ((long)k * L.ny + j) * L.nx + i;
                double xc = x[id], s = 0.0;
                if (i > 0) s += cx * (xc - x[id - 1]);
                if (i < L.nx - 1) s += cx * (xc - x[id + 1]);
                if (j > 0) s += cy * (xc - x[id - L.nx]);
                if (j < L.ny - 1) s += cy * (xc - x[id + L.nx]);
                if (k > 0) s += cz * (xc - x[id - pl]);
                if (k < L.nz - 1) s += cz * (xc - x[id + pl]);
                y[id] = s;
            }
}

void tb_mg_free(tb_ctx *c) {
    for (int l = 0; l < c->nlev; l++) {
        Level &L = c->lev[l];
        if (l > 0) { hipFree(L.x64); hipFree(L.b64); hipFree(L.r64); }
        hipFree(L.x32); hipFree(L.b32); hipFree(L.r32);
        hipFree(L.a64); hipFree(L.a32);
        L.x64 = L.b64 = L.r64 = L.a64 = nullptr;
        L.x32 = L.b32 = L.r32 = L.a32 = nullptr;
    }
    hipFree(c->coarse_inv64); hipFree(c->coarse_inv32);
    c->coarse_inv64 = nullptr; c->coarse_inv32 = nullptr;
    for (int a = 0; a < 2; a++) {
        if (c->vc_exec[a]) hipGraphExecDestroy(c->vc_exec[a]);
        c->vc_exec[a] = nullptr;
    }
    for (int a = 0; a < 3; a++) { hipFree(c->cg32[a]); c->cg32[a] = nullptr; }
    c->nlev = 0;
    c->mg_built = false;
}

int tb_mg_build(tb_ctx *c) {
    if (c->mg_built) return 0;
    Level &L0 = c->lev[0];
    L0 = Level();
    L0.nx = c->ncx; L0.ny = c->ncy; L0.nz = c->ncz;
    L0.px = c->d.px; L0.plane = c->d.plane; L0.ntot = c->d.ntot;
    L0.hx = c->cfg.lx / c->ncx; L0.hy = c->cfg.ly / c->ncy; L0.hz = c->cfg.lz / c->ncz;
    int nl = 1;
    while (nl < 16) {
        Level &L = c->lev[nl - 1];
        if ((long)L.nx * L.ny * L.nz <= COARSEST_CELLS) break;
        int cx, cy, cz;
        if (!choose(L, cx, cy, cz)) break;
        L.cx = cx; L.cy = cy; L.cz = cz;
        Level &C = c->lev[nl];
        C = Level();
        C.nx = cx ? L.nx / 2 : L.nx; C.hx = cx ? 2.0 * L.hx : L.hx;
        C.ny = cy ? L.ny / 2 : L.ny; C.hy = cy ? 2.0 * L.hy : L.hy;
        C.nz = cz ? L.nz / 2 : L.nz; C.hz = cz ? 2.0 * L.hz : L.hz;
        C.px = (C.nx + 15) / 16 * 16;
        C.plane = (long)C.px * C.ny;
        C.ntot = C.plane * C.nz;
        nl++;
    }
    c->nlev = nl;
    for (int l = 0; l < nl; l++) {
        Level &L = c->lev[l];
        if (l > 0) {
            TB_CHECK(c, hipMalloc((void **)&L.x64, sizeof(double) * L.ntot));
            TB_CHECK(c, hipMalloc((void **)&L.b64, sizeof(double) * L.ntot));
            TB_CHECK(c, hipMalloc((void **)&L.r64, sizeof(double) * L.ntot));
            TB_CHECK(c, hipMemsetAsync(L.x64, 0, sizeof(double) * L.ntot, c->s));
            TB_CHECK(c, hipMemsetAsync(L.b64, 0, sizeof(double) * L.ntot, c->s));
            TB_CHECK(c, hipMemsetAsync(L.r64, 0, sizeof(double) * L.ntot, c->s));
            c->bytes += 3 * sizeof(double) * L.ntot;
        }
        // second iterate buffers of the fused sweep pairs (both precisions;
        // with two barriers per plane the FP32 pair beats two half sweeps
        // even though those find their arrays in the 256 MB Infinity Cache)
        L.pairs = (long)L.nx * L.ny * L.nz >= PAIR_LEVEL_CELLS && L.ntot < (1L << 31);
        if (L.pairs) {
            TB_CHECK(c, hipMalloc((void **)&L.a64, sizeof(double) * L.ntot));
            TB_CHECK(c, hipMemsetAsync(L.a64, 0, sizeof(double) * L.ntot, c->s));
            TB_CHECK(c, hipMalloc((void **)&L.a32, sizeof(float) * L.ntot));
            TB_CHECK(c, hipMemsetAsync(L.a32, 0, sizeof(float) * L.ntot, c->s));
            c->bytes += (sizeof(double) + sizeof(float)) * L.ntot;
        }
        TB_CHECK(c, hipMalloc((void **)&L.x32, sizeof(float) * L.ntot));
        TB_CHECK(c, hipMalloc((void **)&L.b32, sizeof(float) * L.ntot));
        TB_CHECK(c, hipMalloc((void **)&L.r32, sizeof(float) * L.ntot));
        TB_CHECK(c, hipMemsetAsync(L.x32, 0, sizeof(float) * L.ntot, c->s));
        TB_CHECK(c, hipMemsetAsync(L.b32, 0, sizeof(float) * L.ntot, c->s));
        TB_CHECK(c, hipMemsetAsync(L.r32, 0, sizeof(float) * L.ntot, c->s));
        c->bytes += 3 * sizeof(float) * L.ntot;
    }
    // vectors of the FP32 inner CG; entries outside the interior cells stay 0
    for (int a = 0; a < 3; a++) {
        TB_CHECK(c, hipMalloc((void **)&c->cg32[a], sizeof(float) * L0.ntot));
        TB_CHECK(c, hipMemsetAsync(c->cg32[a], 0, sizeof(float) * L0.ntot, c->s));
        c->bytes += sizeof(float) * L0.ntot;
    }
    // The coarsest level is solved by its dense inverse when it has at most
    // 1024 cells.  Grid dimensions that stop halving (an odd number of
    // cells) can leave a larger one: it then takes COARSE_SWEEPS symmetric
    // sweep pairs from zero (the oracle's rule) and no LDS cycle is used.
    Level &Lc = c->lev[nl - 1];
    const long nc = (long)Lc.nx * Lc.ny * Lc.nz;
    c->coarse_dense = nc <= 1024;
    c->coarse_n = (int)(nc < (1L << 30) ? nc : (1L << 30));
    // first level of the LDS-resident small cycle, per precision
    for (int a = 0; a < 2 && !c->coarse_dense; a++) c->small_l0[a] = nl;
    for (int a = 0; a < 2 && c->coarse_dense; a++) {
        int l0 = nl - 1;
        while (l0 > 0) {
            const Level &P = c->lev[l0 - 1];
            const int bytes = a ? small_lds_bytes<double>(c, l0 - 1) : small_lds_bytes<float>(c, l0 - 1);
            if ((long)P.nx * P.ny * P.nz > SMALL_LEVEL_CELLS || bytes > SMALL_LDS_BYTES) break;
            l0--;
        }
        c->small_l0[a] = l0;
        const int bytes = a ? small_lds_bytes<double>(c, l0) : small_lds_bytes<float>(c, l0);
        if (bytes > SMALL_LDS_BYTES) { c->err = "multigrid: coarsest levels exceed the LDS budget"; return -1; }
    }
    TB_CHECK(c, hipFuncSetAttribute((const void *)k_small_cycle<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    SMALL_LDS_BYTES));
    TB_CHECK(c, hipFuncSetAttribute((const void *)k_small_cycle<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    SMALL_LDS_BYTES));
    {
        const void *blk[8] = {
            (const void *)k_rbgs_blk<float, 1, true, false, 0>, (const void *)k_rbgs_blk<float, 2, true, false, 0>,
            (const void *)k_rbgs_blk<float, 1, false, true, 1>, (const void *)k_rbgs_blk<float, 2, false, true, 1>,
            (const void *)k_rbgs_blk<double, 1, true, false, 0>, (const void *)k_rbgs_blk<double, 2, true, false, 0>,
            (const void *)k_rbgs_blk<double, 1, false, true, 1>, (const void *)k_rbgs_blk<double, 2, false, true, 1>};
        for (int a = 0; a < 8; a++)
            TB_CHECK(c, hipFuncSetAttribute(blk[a], hipFuncAttributeMaxDynamicSharedMemorySize, BLK_LDS_BYTES));
    }
    if (!c->coarse_dense) {
        TB_CHECK(c, hipStreamSynchronize(c->s));
        c->mg_built = true;
        return 0;
    }
    // dense inverse of (A_c + J/n) on the coarsest level (Gauss-Jordan,
    // partial pivoting), FP64 on the host
    const int n = (int)nc;
    std::vector<double> A((size_t)n * n), I((size_t)n * n, 0.0), e(n), col(n);
    for (int a = 0; a < n; a++) {
        std::fill(e.begin(), e.end(), 0.0);
        e[a] = 1.0;
        level_apply_host(Lc, e.data(), col.data());
        for (int r = 0; r < n; r++) A[(size_t)r * n + a] = col[r] + 1.0 / n;
        I[(size_t)a * n + a] = 1.0;
    }
    for (int cc = 0; cc < n; cc++) {
        int piv = cc;
        for (int r = cc + 1; r < n; r++) if (fabs(A[(size_t)r * n + cc]) > fabs(A[(size_t)piv * n + cc])) piv = r;
        if (piv != cc)
            for (int k = 0; k < n; k++) {
                std::swap(A[(size_t)cc * n + k], A[(size_t)piv * n + k]);
                std::swap(I[(size_t)cc * n + k], I[(size_t)piv * n + k]);
            }
        double dd = 1.0 / A[(size_t)cc * n + cc];
        for (int k = 0; k < n; k++) { A[(size_t)cc * n + k] *= dd; I[(size_t)cc * n + k] *= dd; }
        for (int r = 0; r < n; r++) {
            if (r == cc) continue;
            double f = A[(size_t)r * n + cc];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) { A[(size_t)r * n + k] -= f * A[(size_t)cc * n + k]; I[(size_t)r * n + k] -= f * I[(size_t)cc * n + k]; }
        }
    }
    // stored transposed (column cc contiguous) so that the threads of the
    // coarse solve, one per row, read consecutive addresses
    std::vector<double> IT((size_t)n * n);
    for (int r = 0; r < n; r++)
        for (int cc = 0; cc < n; cc++) IT[(size_t)cc * n + r] = I[(size_t)r * n + cc];
    I.swap(IT);
    std::vector<float> I32((size_t)n * n);
    for (size_t a = 0; a < I.size(); a++) I32[a] = (float)I[a];
    TB_CHECK(c, hipMalloc((void **)&c->coarse_inv64, sizeof(double) * (size_t)n * n));
    TB_CHECK(c, hipMalloc((void **)&c->coarse_inv32, sizeof(float) * (size_t)n * n));
    TB_CHECK(c, hipMemcpy(c->coarse_inv64, I.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice));
    TB_CHECK(c, hipMemcpy(c->coarse_inv32, I32.data(), sizeof(float) * (size_t)n * n, hipMemcpyHostToDevice));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    c->mg_built = true;
    return 0;
}

template <typename T>
struct MGBuf {
    T *x, *b, *r, *a;   // a: second iterate buffer of the fused-pair levels, else null
};

template <typename T>
static MGBuf<T> bufs(tb_ctx *c, int l);
template <>
MGBuf<double> bufs<double>(tb_ctx *c, int l) { return {c->lev[l].x64, c->lev[l].b64, c->lev[l].r64, c->lev[l].a64}; }
template <>
MGBuf<float> bufs<float>(tb_ctx *c, int l) { return {c->lev[l].x32, c->lev[l].b32, c->lev[l].r32, c->lev[l].a32}; }

template <typename T>
static const T *coarse_inv(tb_ctx *c);
template <>
const double *coarse_inv<double>(tb_ctx *c) { return c->coarse_inv64; }
template <>
const float *coarse_inv<float>(tb_ctx *c) { return c->coarse_inv32; }

// the column-walking transfer kernels serve large levels coarsened in y and
// z (on small ones their plane-by-plane chains cost more than two launches)
constexpr long FUSED_TRANSFER_CELLS = 1L << 20;
static bool transfer_011(const Level &L) {
    return L.cx == 0 && L.cy == 1 && L.cz == 1 && L.ntot < (1L << 31) &&
           (long)L.nx * L.ny * L.nz >= FUSED_TRANSFER_CELLS;
}

// x_f = x_in + P x_c from level l+1 to level l
template <typename T>
static void prolong(tb_ctx *c, int l, const T *xc, const T *xin, T *xf) {
    const Level &L = c->lev[l];
    LV<T> lv = lv_of<T>(L);
    const Level &Cl = c->lev[l + 1];
    LV<T> cv = lv_of<T>(Cl);
    if (transfer_011(L)) {
        KTiles t;
        t.nx = L.nx; t.ny = L.ny; t.nz = L.nz;
        t.tx = (L.nx + TILE_X - 1) / TILE_X;
        t.ty = (L.ny + TILE_Y - 1) / TILE_Y;
        t.kc = (L.nz + PR_KC - 1) / PR_KC;
        t.n = (long)t.tx * t.ty * t.kc;
        t.per_xcd = (t.n + NXCD - 1) / NXCD;
        hipLaunchKernelGGL(k_prolong_011<T>, ktiles_grid(t), tiles_block(), 0, c->s, lv, cv, xc, xin, xf, t);
        return;
    }
    Tiles all = make_tiles(L.nx, L.ny, L.nz);
    hipLaunchKernelGGL(k_prolong<T>, tiles_grid(all), tiles_block(), 0, c->s, lv, cv, L.cx, L.cy, L.cz, xc, xin, xf,
                       all);
}

// b of level l+1 from the residual of x on level l (fused for (0, 1, 1))
template <typename T>
static void resid_restrict(tb_ctx *c, int l, const T *x) {
    const Level &L = c->lev[l];
    MGBuf<T> B = bufs<T>(c, l);
    LV<T> lv = lv_of<T>(L);
    const Level &Cl = c->lev[l + 1];
    MGBuf<T> CB = bufs<T>(c, l + 1);
    LV<T> cv = lv_of<T>(Cl);
    if (transfer_011(L)) {
        KTiles t;
        t.nx = Cl.nx; t.ny = Cl.ny; t.nz = Cl.nz;
        t.tx = (Cl.nx + TILE_X - 1) / TILE_X;
        t.ty = (Cl.ny + TILE_Y - 1) / TILE_Y;
        t.kc = (Cl.nz + RR_KC - 1) / RR_KC;
        t.n = (long)t.tx * t.ty * t.kc;
        t.per_xcd = (t.n + NXCD - 1) / NXCD;
        hipLaunchKernelGGL(k_resid_restrict_011<T>, ktiles_grid(t), tiles_block(), 0, c->s, lv, cv, x,
                           (const T *)B.b, CB.b, t);
        return;
    }
    Tiles all = make_tiles(L.nx, L.ny, L.nz);
    hipLaunchKernelGGL(k_mg_resid<T>, tiles_grid(all), tiles_block(), 0, c->s, lv, x, (const T *)B.b, B.r, all);
    Tiles call = make_tiles(Cl.nx, Cl.ny, Cl.nz);
    const T *r = B.r;
    const dim3 g = tiles_grid(call), blk = tiles_block();
    switch (L.cx * 4 + L.cy * 2 + L.cz) {
    case 1: hipLaunchKernelGGL((k_restrict_t<T, 0, 0, 1>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 2: hipLaunchKernelGGL((k_restrict_t<T, 0, 1, 0>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 3: hipLaunchKernelGGL((k_restrict_t<T, 0, 1, 1>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 4: hipLaunchKernelGGL((k_restrict_t<T, 1, 0, 0>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 5: hipLaunchKernelGGL((k_restrict_t<T, 1, 0, 1>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 6: hipLaunchKernelGGL((k_restrict_t<T, 1, 1, 0>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    case 7: hipLaunchKernelGGL((k_restrict_t<T, 1, 1, 1>), g, blk, 0, c->s, lv, cv, r, CB.b, call); break;
    default:
        hipLaunchKernelGGL(k_restrict<T>, g, blk, 0, c->s, lv, cv, L.cx, L.cy, L.cz, r, CB.b, call);
        break;
    }
}

// symmetric sweep pairs from zero on a coarsest level too large for the
// dense inverse (mg_build): forward pairs, then as many backward ones
constexpr int COARSE_SWEEPS = 50;

// the two kinds of blocked launch: pre-smoothing from zero, prolongation +
// post-smoothing
enum { BLK_ZERO = 0, BLK_PROL };

template <typename T, int NP>
static void launch_blk_np(tb_ctx *c, const dim3 &grid, int lds, int kind, const LV<T> &lv, const T *src,
                          const T *b, T *dst, const BlkShape &S, const LV<T> &cv, const Level &L, const T *xc) {
    const dim3 blk(BLK_THREADS);
    if (kind == BLK_ZERO)
        hipLaunchKernelGGL((k_rbgs_blk<T, NP, true, false, 0>), grid, blk, lds, c->s, lv, src, b, dst, S, cv, L.cx,
                           L.cy, L.cz, xc);
    else
        hipLaunchKernelGGL((k_rbgs_blk<T, NP, false, true, 1>), grid, blk, lds, c->s, lv, src, b, dst, S, cv, L.cx,
                           L.cy, L.cz, xc);
}

template <typename T>
static void launch_blk(tb_ctx *c, int l, int np, int kind, const T *src, T *dst, const T *xc) {
    const Level &L = c->lev[l];
    MGBuf<T> B = bufs<T>(c, l);
    LV<T> lv = lv_of<T>(L);
    const Level &Cl = c->lev[l + 1];
    LV<T> cv = lv_of<T>(Cl);
    BlkShape S;
    int lds = 0;
    blk_shape(L, np, sizeof(T), S, lds);   // checked by blk_usable
    const dim3 grid((unsigned)(S.per_xcd * NXCD));
    if (np == 1) launch_blk_np<T, 1>(c, grid, lds, kind, lv, src, (const T *)B.b, dst, S, cv, L, xc);
    else launch_blk_np<T, 2>(c, grid, lds, kind, lv, src, (const T *)B.b, dst, S, cv, L, xc);
}

// the blocked smoother serves a level below the finest without fused pairs
// that has at most BLK_MAX_CELLS cells, with 1 or 2 sweep pairs per side and
// a box whose LDS fits
static bool blk_usable(const tb_ctx *c, int l, int pre, int post, size_t tsize) {
    if (!c->mg_blocked || l == 0 || pre < 1 || pre > 2 || post < 1 || post > 2 || c->lev[l].pairs) return false;
    if ((long)c->lev[l].nx * c->lev[l].ny * c->lev[l].nz > BLK_MAX_CELLS) return false;
    BlkShape S;
    int lds = 0;
    return blk_shape(c->lev[l], pre, tsize, S, lds) && blk_shape(c->lev[l], post, tsize, S, lds);
}

// One V-cycle from level l down; *res receives the buffer that holds the
// level's result (x, or r for the blocked smoother, whose post-smoother
// cannot overwrite the x its neighbours' halos read).
template <typename T>
static int vcycle_level(tb_ctx *c, int l, const T **res) {
    const Level &L = c->lev[l];
    MGBuf<T> B = bufs<T>(c, l);
    LV<T> lv = lv_of<T>(L);
    if (l == 0) c->vc_dot_blocks[sizeof(T) == 8 ? 1 : 0] = 0;
    if (l == c->nlev - 1 && !c->coarse_dense) {
        const Tiles half = make_tiles((L.nx + 1) / 2, L.ny, L.nz);
        auto sweep = [&](int color, int zero_init) {
            TB_PROF(c, TB_K_MG_SMALL,
                    hipLaunchKernelGGL(k_rbgs<T>, tiles_grid(half), tiles_block(), 0, c->s, lv, B.x, (const T *)B.b,
                                       color, zero_init, half));
        };
        sweep(0, 1);
        sweep(1, 0);
        for (int s = 1; s < COARSE_SWEEPS; s++) {
            sweep(0, 0);
            sweep(1, 0);
        }
        for (int s = 0; s < COARSE_SWEEPS; s++) {
            sweep(1, 0);
            sweep(0, 0);
        }
        TB_LAUNCH_CHECK(c);
        *res = B.x;
        return 0;
    }
    if (l == c->small_l0[sizeof(T) == 8 ? 1 : 0]) {
        SmallCycle<T> A;
        A.l0 = l;
        A.nl = c->nlev;
        A.pre = c->cfg.mg_pre;
        A.post = c->cfg.mg_post;
        A.coarse_n = c->coarse_n;
        A.Minv = coarse_inv<T>(c);
        A.g0 = lv;
        A.gx0 = B.x;
        A.gb0 = B.b;
        int off = 0;
        for (int m = l; m < c->nlev; m++) {
            const Level &Lm = c->lev[m];
            LV<T> v = lv_of<T>(Lm);
            v.px = Lm.nx;
            v.py = Lm.ny;
            A.lv[m] = v;
            A.xo[m] = off; off += Lm.nx * Lm.ny * Lm.nz;
            A.bo[m] = off; off += Lm.nx * Lm.ny * Lm.nz;
            A.cx[m] = Lm.cx; A.cy[m] = Lm.cy; A.cz[m] = Lm.cz;
        }
        A.ro = off;
        TB_PROF(c, TB_K_MG_SMALL,
                hipLaunchKernelGGL(k_small_cycle<T>, dim3(1), dim3(1024), small_lds_bytes<T>(c, l), c->s, A));
        TB_LAUNCH_CHECK(c);
        *res = B.x;
        return 0;
    }
    Tiles all = make_tiles(L.nx, L.ny, L.nz);
    Tiles half = make_tiles((L.nx + 1) / 2, L.ny, L.nz);
    const int kid = (l == 0) ? (sizeof(T) == 8 ? TB_K_RBGS0 : TB_K_RBGS0_F32) : TB_K_MG_LEVELS;
    const int prec = sizeof(T) == 8 ? 1 : 0;
    const int pre = (l == 0 && c->mg_pre0[prec] > 0) ? c->mg_pre0[prec] : c->cfg.mg_pre;
    const int post = (l == 0 && c->mg_post0[prec] > 0) ? c->mg_post0[prec] : c->cfg.mg_post;
    // large levels: fused sweep pairs alternating between x and the level's
    // second buffer; cur is the buffer holding the current iterate
    const bool pairs = L.pairs;
    T *cur = B.x;
    if (pairs) {
        const PairTiles P = make_pair_tiles(L.nx, L.ny, L.nz, pair_chunk(L));
        const dim3 grid(P.per_xcd * NXCD), blk(PAIR_THREADS);
        // the first post pair forms the prolongation itself on (0, 1, 1)
        // levels; the pairs then write pre + post times, so the first one
        // goes to the buffer that makes the last one land in x
        const bool fuse = sizeof(T) == 4 && c->mg_blocked && transfer_011(L) && pre > 0 && post > 0;
        T *zdst = (fuse && (pre + post) % 2 == 0) ? B.a : B.x;
        const Level &Cl = c->lev[l + 1];
        const LV<T> cv = lv_of<T>(Cl);
        auto pair = [&](int first, bool zero, bool dot, const T *pxc) {
            T *dst = zero ? zdst : (cur == B.x ? B.a : B.x);
            const T *src = cur;
            const T *bb = B.b;
            if (c->prof_on && kid >= 0) tb_prof_begin(c, kid);
            if (zero) hipLaunchKernelGGL((k_rbgs_pair<T, 0, true, false>), grid, blk, 0, c->s, lv, src, dst, bb, P, nullptr, lv, nullptr);
            else if (first == 0) hipLaunchKernelGGL((k_rbgs_pair<T, 0, false, false>), grid, blk, 0, c->s, lv, src, dst, bb, P, nullptr, lv, nullptr);
            else if (pxc) {
                // FP32 only: the FP64 form needs more registers than two blocks per CU leave
                if constexpr (sizeof(T) == 4) {
                    if (!dot) hipLaunchKernelGGL((k_rbgs_pair<T, 1, false, false, true>), grid, blk, 0, c->s, lv, src, dst, bb, P, nullptr, cv, pxc);
                    else hipLaunchKernelGGL((k_rbgs_pair<T, 1, false, true, true>), grid, blk, 0, c->s, lv, src, dst, bb, P, c->red, cv, pxc);
                }
            } else if (!dot) hipLaunchKernelGGL((k_rbgs_pair<T, 1, false, false>), grid, blk, 0, c->s, lv, src, dst, bb, P, nullptr, lv, nullptr);
            else hipLaunchKernelGGL((k_rbgs_pair<T, 1, false, true>), grid, blk, 0, c->s, lv, src, dst, bb, P, c->red, lv, nullptr);
            if (c->prof_on && kid >= 0) tb_prof_end(c, kid);
            cur = dst;
        };
        if (pre > 0) {
            pair(0, true, false, nullptr);
            for (int s = 1; s < pre; s++) pair(0, false, false, nullptr);
        } else {
            hipLaunchKernelGGL(k_zero<T>, tiles_grid(all), tiles_block(), 0, c->s, lv, B.x, all);
        }
        TB_PROF(c, TB_K_MG_TRANSFER, resid_restrict<T>(c, l, (const T *)cur));
        TB_LAUNCH_CHECK(c);
        const T *xc = nullptr;
        if (vcycle_level<T>(c, l + 1, &xc)) return -1;
        if (!fuse) {
            // the post pairs alternate buffers, so the prolonged iterate goes
            // to the buffer from which they end in x
            T *dst = (post % 2 == 0) ? B.x : B.a;
            TB_PROF(c, TB_K_MG_TRANSFER, prolong<T>(c, l, xc, (const T *)cur, dst));
            cur = dst;
        }
        // on the finest level the last pair also forms the partial sums of
        // r . z (b . x) for the Krylov method around the cycle
        const bool dot = l == 0 && post > 0;
        for (int s = 0; s < post; s++) pair(1, false, dot && s == post - 1, (fuse && s == 0) ? xc : nullptr);
        if (l == 0) {
            c->vc_dot_blocks[sizeof(T) == 8 ? 1 : 0] = dot ? (int)(P.per_xcd * NXCD) : 0;
        }
        TB_LAUNCH_CHECK(c);
        *res = B.x;
        return 0;
    }
    if (blk_usable(c, l, pre, post, sizeof(T))) {
        // pre-smoother from x = 0 into x, residual and restriction, the
        // coarse levels, then prolongation + post-smoother from x into r
        TB_PROF(c, kid, launch_blk<T>(c, l, pre, BLK_ZERO, nullptr, B.x, nullptr));
        TB_PROF(c, TB_K_MG_TRANSFER, resid_restrict<T>(c, l, (const T *)B.x));
        TB_LAUNCH_CHECK(c);
        const T *xc = nullptr;
        if (vcycle_level<T>(c, l + 1, &xc)) return -1;
        TB_PROF(c, kid, launch_blk<T>(c, l, post, BLK_PROL, (const T *)B.x, B.r, xc));
        TB_LAUNCH_CHECK(c);
        *res = B.r;
        return 0;
    }
    auto smooth = [&](int color, int zero_init) {
        if (c->prof_on && kid >= 0) tb_prof_begin(c, kid);
        hipLaunchKernelGGL(k_rbgs<T>, tiles_grid(half), tiles_block(), 0, c->s, lv, B.x, (const T *)B.b, color,
                           zero_init, half);
        if (c->prof_on && kid >= 0) tb_prof_end(c, kid);
    };
    // zero initial guess: the first red half sweep reads no neighbour and
    // the black one only red cells, so x needs no clearing pass
    if (pre > 0) {
        smooth(0, 1);
        smooth(1, 0);
        for (int s = 1; s < pre; s++) {
            smooth(0, 0);
            smooth(1, 0);
        }
    } else {
        hipLaunchKernelGGL(k_zero<T>, tiles_grid(all), tiles_block(), 0, c->s, lv, B.x, all);
    }
    TB_PROF(c, TB_K_MG_TRANSFER, resid_restrict<T>(c, l, (const T *)B.x));
    TB_LAUNCH_CHECK(c);
    const T *xc = nullptr;
    if (vcycle_level<T>(c, l + 1, &xc)) return -1;
    TB_PROF(c, TB_K_MG_TRANSFER, prolong<T>(c, l, xc, (const T *)B.x, B.x));
    for (int s = 0; s < post; s++) {
        smooth(1, 0);
        smooth(0, 0);
    }
    TB_LAUNCH_CHECK(c);
    *res = B.x;
    return 0;
}

// z = V(r) in FP64 with level 0 aliased onto the padded vectors
// A V-cycle is ~10 levels x ~12 launches, most of them on grids far too
// small to fill the chip, so it is captured once into a HIP graph and
// replayed with one launch.  The graph holds kernels only (no memset nodes).
// Profiling runs launch the kernels directly, so their events are recorded.
template <typename T>
static int vcycle_run(tb_ctx *c, int which) {
    const T *res = nullptr;
    if (!c->use_graphs || c->prof_on) return vcycle_level<T>(c, 0, &res);
    if (!c->vc_exec[which]) {
        hipGraph_t g = nullptr;
        TB_CHECK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        const int rc = vcycle_level<T>(c, 0, &res);
        const hipError_t e = hipStreamEndCapture(c->s, &g);
        if (rc || e != hipSuccess) {
            if (e == hipSuccess && g) hipGraphDestroy(g);
            if (!rc) c->err = std::string("V-cycle capture: ") + hipGetErrorString(e);
            return -1;
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) { c->err = std::string("V-cycle graph: ") + hipGetErrorString(ei); return -1; }
        c->vc_exec[which] = ex;
    }
    TB_CHECK(c, hipGraphLaunch(c->vc_exec[which], c->s));
    return 0;
}

int tb_vcycle64(tb_ctx *c, const double *r_pad, double *z_pad) {
    if (tb_mg_build(c)) return -1;
    const bool same = c->lev[0].b64 == r_pad && c->lev[0].x64 == z_pad && c->lev[0].r64 == c->kv[4];
    if (!same && c->vc_exec[0]) {        // captured with other level-0 vectors
        hipGraphExecDestroy(c->vc_exec[0]);
        c->vc_exec[0] = nullptr;
    }
    c->lev[0].b64 = const_cast<double *>(r_pad);
    c->lev[0].x64 = z_pad;
    c->lev[0].r64 = c->kv[4];
    return vcycle_run<double>(c, 0);
}

// lev[0].x32 = V(lev[0].b32) in FP32
int tb_vcycle32(tb_ctx *c) {
    if (tb_mg_build(c)) return -1;
    return vcycle_run<float>(c, 1);
}

// sweeps (or sweep pairs) on the finest level of the FP32 / FP64 V-cycles;
// 0 keeps cfg.mg_pre / cfg.mg_post there.  Captured V-cycles are dropped.
extern "C" int tb_set_mg_finest_sweeps(tb_ctx *c, int pre32, int post32, int pre64, int post64) {
    if (!c) return -1;
    if (pre32 < 0 || post32 < 0 || pre64 < 0 || post64 < 0 || pre32 > 8 || post32 > 8 || pre64 > 8 || post64 > 8) {
        c->err = "tb_set_mg_finest_sweeps: sweeps must be 0..8";
        return -1;
    }
    c->mg_pre0[0] = pre32;
    c->mg_post0[0] = post32;
    c->mg_pre0[1] = pre64;
    c->mg_post0[1] = post64;
    for (int a = 0; a < 2; a++) {
        if (c->vc_exec[a]) hipGraphExecDestroy(c->vc_exec[a]);
        c->vc_exec[a] = nullptr;
    }
    return 0;
}

extern "C" int tb_set_mg_smoother(tb_ctx *c, int blocked) {
    if (!c) return -1;
    if (blocked != 0 && blocked != 1) {
        c->err = "tb_set_mg_smoother: blocked must be 0 or 1";
        return -1;
    }
    c->mg_blocked = blocked != 0;
    for (int a = 0; a < 2; a++) {
        if (c->vc_exec[a]) hipGraphExecDestroy(c->vc_exec[a]);
        c->vc_exec[a] = nullptr;
    }
    return 0;
}

extern "C" int tb_set_graphs(tb_ctx *c, int on) {
    if (!c) return -1;
    c->use_graphs = on != 0;
    return 0;
}

// ---------------------------------------------------------------------------
// vector kernels of the Poisson solvers (grid-stride over the padded array)
constexpr int PB = 1024;
constexpr int PT = 256;

// x += alpha p ; r -= alpha q ; partial r.r      (alpha = rz / pq)
__global__ void k_pcg_xr(long n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                         const double *__restrict__ q, const double *__restrict__ s, double *part) {
    __shared__ double sh[16];
    double alpha = s[S_RZ] / s[S_PQ];
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// p = z + beta p with beta = rz_new / rz ; rz <- rz_new (thread 0 of block 0
// writes it after every block has read the old value: done by the next launch)
__global__ void k_pcg_p(long n, const double *__restrict__ z, double *__restrict__ p, const double *__restrict__ s) {
    double beta = s[S_TMP0] / s[S_RZ];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

__global__ void k_copy_scalar(double *s, int dst, int src) {
    if (threadIdx.x == 0) s[dst] = s[src];
}

// x -= x[corner] on non-boundary cells ; boundary entries set to 0
__global__ void __launch_bounds__(256) k_shift(Dims d, double *__restrict__ x, const double *__restrict__ s, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    bool bnd = (i == d.NX - 1) || (j == d.NY - 1) || (k == d.NZ - 1);
    x[id] = bnd ? 0.0 : x[id] - s[S_CORNER];
}

// the corner value after x += e: s[S_CORNER] = x[corner] + (double) e[corner]
__global__ void k_corner_after_add(const double *__restrict__ x, const float *__restrict__ e, long corner,
                                   double *__restrict__ s) {
    if (threadIdx.x == 0) s[S_CORNER] = x[corner] + (double)e[corner];
}

// x += (double) e32 followed by the pinned shift, in one pass over the cells:
// x = (x + e) - s[S_CORNER] (the same values as adding e and then
// k_shift; the boundary entries stay 0)
__global__ void __launch_bounds__(256) k_add32_pin(Dims d, double *__restrict__ x, const float *__restrict__ e,
                                                   const double *__restrict__ s, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    x[id] = (x[id] + (double)e[id]) - s[S_CORNER];
}

static int read_s(tb_ctx *c, int first, int count) { return tb_read_scalars(c, first, count); }

struct PoisCtx {
    Tiles all, cells;
    KTiles kt;
    PoisOp Pp, Pn;
    long nblk;
    long corner;
};

static PoisCtx pois_ctx(tb_ctx *c) {
    PoisCtx P;
    P.all = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    P.cells = make_tiles(c->ncx, c->ncy, c->ncz);
    P.Pp = make_pois(c, 0);
    P.Pn = make_pois(c, 1);
    P.kt = make_ktiles(c->d.NX, c->d.NY, c->d.NZ);
    P.nblk = P.kt.per_xcd * NXCD;
    P.corner = gidx(c->d, c->ncx - 1, c->ncy - 1, c->ncz - 1);
    return P;
}

// shift so the pinned corner holds 0
static int shift_corner(tb_ctx *c, const PoisCtx &P, double *x) {
    TB_CHECK(c, hipMemcpyAsync(c->scal + S_CORNER, x + P.corner, sizeof(double), hipMemcpyDeviceToDevice, c->s));
    hipLaunchKernelGGL(k_shift, tiles_grid(P.all), tiles_block(), 0, c->s, c->d, x, (const double *)c->scal, P.all);
    TB_LAUNCH_CHECK(c);
    return 0;
}

// r = b - A x (pinned or Neumann), returns ||r||^2 and ||b||^2 via hscal
static int residual(tb_ctx *c, const PoisCtx &P, bool neumann, const double *x, const double *b, double *r,
                    float *r32, double &rr, double &bb) {
    hipLaunchKernelGGL(k_pois_km<1>, ktiles_grid(P.kt), tiles_block(), 0, c->s, neumann ? P.Pn : P.Pp, x, b, r, r32,
                       c->red, P.kt);
    TB_LAUNCH_CHECK(c);
    if (tb_reduce_finish(c, (int)P.nblk, 3, S_RES)) return -1;
    if (read_s(c, S_RES, 3)) return -1;
    rr = c->hscal[S_RES];
    bb = c->hscal[S_RES + 1];   // x.b in S_RES + 2
    return 0;
}

// FP64 GMG-preconditioned CG from the current x; returns iterations
// rr_known: x is already pinned and kv[0] holds its residual, of squared norm
// *rr_known; the first shift and residual are skipped.
static int pcg64(tb_ctx *c, const PoisCtx &P, double *x, const double *b, double bn, double rtol, int maxit,
                 int &its, int &vcy, double &res, const double *rr_known) {
    long n = c->d.ntot;
    double *r = c->kv[0], *z = c->kv[1], *p = c->kv[2], *q = c->kv[3];
    double rr, bb;
    // the V-cycle writes interior cells only; keep z's half-cell planes at 0
    TB_CHECK(c, hipMemsetAsync(z, 0, sizeof(double) * n, c->s));
    if (rr_known) {
        rr = *rr_known;
    } else {
        if (shift_corner(c, P, x)) return -1;
        if (residual(c, P, false, x, b, r, nullptr, rr, bb)) return -1;
    }
    res = sqrt(rr) / bn;
    // r . z of the V-cycle just run: from its last pair's partial sums, or a
    // separate pass when the finest level has no pairs
    auto rz = [&](int slot) -> int {
        if (c->vc_dot_blocks[1] > 0) return tb_reduce_finish(c, c->vc_dot_blocks[1], 1, slot);
        return tb_dot(c, r, z, slot);
    };
    while (res > rtol && its < maxit) {
        if (tb_vcycle64(c, r, z)) return -1;
        vcy++;
        TB_CHECK(c, hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s));
        if (rz(S_RZ)) return -1;
        int inner = 0;
        while (its < maxit) {
            TB_PROF(c, TB_K_POIS_OP,
                    hipLaunchKernelGGL(k_pois_km<0>, ktiles_grid(P.kt), tiles_block(), 0, c->s, P.Pn,
                                       (const double *)p, (const double *)nullptr, q, (float *)nullptr, c->red, P.kt));
            TB_LAUNCH_CHECK(c);
            if (tb_reduce_finish(c, (int)P.nblk, 1, S_PQ)) return -1;
            TB_PROF(c, TB_K_PCG_UPD,
                    hipLaunchKernelGGL(k_pcg_xr, dim3(PB), dim3(PT), 0, c->s, n, x, r, (const double *)p,
                                       (const double *)q, (const double *)c->scal, c->red));
            TB_LAUNCH_CHECK(c);
            if (tb_reduce_finish(c, PB, 1, S_RR)) return -1;
            its++;
            inner++;
            if (read_s(c, S_RR, 1)) return -1;
            double rrel = sqrt(c->hscal[S_RR]) / bn;
            if (!(rrel == rrel)) { c->err = "Poisson CG produced NaN"; return -1; }
            if (rrel <= 0.5 * rtol || inner >= 200) break;
            if (tb_vcycle64(c, r, z)) return -1;
            vcy++;
            if (rz(S_TMP0)) return -1;
            hipLaunchKernelGGL(k_pcg_p, dim3(PB), dim3(PT), 0, c->s, n, (const double *)z, p, (const double *)c->scal);
            hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(64), 0, c->s, c->scal, S_RZ, S_TMP0);
            TB_LAUNCH_CHECK(c);
        }
        if (shift_corner(c, P, x)) return -1;
        if (residual(c, P, false, x, b, r, nullptr, rr, bb)) return -1;
        res = sqrt(rr) / bn;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// FP32 inner CG of the mixed-precision solve.  Vectors live in the level-0
// layout (the padded field layout); every entry outside the interior cells is
// 0 and stays 0, so the vector kernels sweep the whole array as float4.  Dot
// products accumulate in FP64.

// q = A p (Neumann -Laplacian of level 0) with the partial p.q
__global__ void __launch_bounds__(256) k_cg32_apply(LV<float> L, const float *__restrict__ p, float *__restrict__ q,
                                                    double *part, KTiles t) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    int i, j, k0, k1;
    bool act;
    ktile_column(t, i, j, k0, k1, act);
    if (act) {
        const long pl = (long)L.px * L.py;
        // absent neighbours are clamped to the cell itself and contribute 0
        const long ow = i > 0 ? 1 : 0, oe = i < L.nx - 1 ? 1 : 0;
        const long os = j > 0 ? L.px : 0, on = j < L.ny - 1 ? L.px : 0;
        long id = lidx(L, i, j, k0);
        float pm = p[id - (k0 > 0 ? pl : 0)];
        float pc = p[id];
        for (int k = k0; k < k1; k++) {
            const float pp = p[id + (k < L.nz - 1 ? pl : 0)];
            const float pw = p[id - ow], pe = p[id + oe], ps = p[id - os], pn = p[id + on];
            const float v = L.cx * (pc - pw) + L.cx * (pc - pe) + L.cy * (pc - ps) + L.cy * (pc - pn) +
                            L.cz * (pc - pm) + L.cz * (pc - pp);
            st_stream(&q[id], v);
            acc[0] += (double)pc * (double)v;
            pm = pc;
            pc = pp;
            id += pl;
        }
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x] = acc[0];
}

// e += alpha p ; r -= alpha q ; partial r.r      (alpha = rz / pq)
__global__ void __launch_bounds__(PT) k_cg32_xr(long n4, float4 *__restrict__ e, float4 *__restrict__ r,
                                                const float4 *__restrict__ p, const float4 *__restrict__ q,
                                                const double *__restrict__ s, double *part) {
    __shared__ double sh[16];
    const float alpha = (float)(s[S_RZ] / s[S_PQ]);
    double acc[1] = {0.0};
    for (long a = blockIdx.x * (long)blockDim.x + threadIdx.x; a < n4; a += (long)gridDim.x * blockDim.x) {
        const float4 pv = p[a], qv = q[a];
        float4 ev = e[a], rv = r[a];
        ev.x += alpha * pv.x; ev.y += alpha * pv.y; ev.z += alpha * pv.z; ev.w += alpha * pv.w;
        rv.x -= alpha * qv.x; rv.y -= alpha * qv.y; rv.z -= alpha * qv.z; rv.w -= alpha * qv.w;
        e[a] = ev;
        r[a] = rv;
        acc[0] += (double)rv.x * rv.x + (double)rv.y * rv.y + (double)rv.z * rv.z + (double)rv.w * rv.w;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(PT) k_dot32(long n4, const float4 *__restrict__ a, const float4 *__restrict__ b,
                                              double *part) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
        const float4 u = a[i], v = b[i];
        acc[0] += (double)u.x * v.x + (double)u.y * v.y + (double)u.z * v.z + (double)u.w * v.w;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// p = z + beta p, beta = rz_new / rz
__global__ void __launch_bounds__(PT) k_cg32_p(long n4, const float4 *__restrict__ z, float4 *__restrict__ p,
                                               const double *__restrict__ s) {
    const float beta = (float)(s[S_TMP0] / s[S_RZ]);
    for (long a = blockIdx.x * (long)blockDim.x + threadIdx.x; a < n4; a += (long)gridDim.x * blockDim.x) {
        const float4 zv = z[a];
        float4 pv = p[a];
        pv.x = zv.x + beta * pv.x; pv.y = zv.y + beta * pv.y; pv.z = zv.z + beta * pv.z; pv.w = zv.w + beta * pv.w;
        p[a] = pv;
    }
}

// Inner solve A e = r on level 0 by FP32 CG preconditioned with FP32
// V-cycles, from e = 0, until ||r|| <= tol_rel * ||r0|| or maxit iterations.
// r is lev[0].b32 (updated in place), z is lev[0].x32, e is cg32[2].
// The first `skip` iterations run without reading their residual back (each
// read-back stalls the stream for a host round trip); the caller takes skip
// from the iterations an earlier solve needed.
static int cg32_inner(tb_ctx *c, const KTiles &kt, double r0n, double tol_rel, int maxit, int skip, int &its,
                      int &vcy) {
    Level &L0 = c->lev[0];
    const long n4 = L0.ntot / 4;
    float *r = L0.b32, *z = L0.x32, *p = c->cg32[0], *q = c->cg32[1], *e = c->cg32[2];
    const LV<float> lv = lv_of<float>(L0);
    const int nblk = (int)(kt.per_xcd * NXCD);
    TB_CHECK(c, hipMemsetAsync(e, 0, sizeof(float) * L0.ntot, c->s));
    // r . z of the V-cycle just run: from its last pair's partial sums, or
    // a separate pass when the finest level has no pairs
    auto rz = [&](int slot) -> int {
        if (c->vc_dot_blocks[0] > 0) return tb_reduce_finish(c, c->vc_dot_blocks[0], 1, slot);
        hipLaunchKernelGGL(k_dot32, dim3(PB), dim3(PT), 0, c->s, n4, (const float4 *)r, (const float4 *)z, c->red);
        TB_LAUNCH_CHECK(c);
        return tb_reduce_finish(c, PB, 1, slot);
    };
    if (tb_vcycle32(c)) return -1;
    vcy++;
    TB_CHECK(c, hipMemcpyAsync(p, z, sizeof(float) * L0.ntot, hipMemcpyDeviceToDevice, c->s));
    if (rz(S_RZ)) return -1;
    int inner = 0;
    while (inner < maxit) {
        TB_PROF(c, TB_K_CG32_OP,
                hipLaunchKernelGGL(k_cg32_apply, ktiles_grid(kt), tiles_block(), 0, c->s, lv, (const float *)p, q,
                                   c->red, kt));
        TB_LAUNCH_CHECK(c);
        if (tb_reduce_finish(c, nblk, 1, S_PQ)) return -1;
        TB_PROF(c, TB_K_CG32_UPD,
                hipLaunchKernelGGL(k_cg32_xr, dim3(PB), dim3(PT), 0, c->s, n4, (float4 *)e, (float4 *)r,
                                   (const float4 *)p, (const float4 *)q, (const double *)c->scal, c->red));
        TB_LAUNCH_CHECK(c);
        if (tb_reduce_finish(c, PB, 1, S_RR)) return -1;
        inner++;
        its++;
        if (inner >= maxit) break;
        if (inner > skip) {
            if (read_s(c, S_RR, 1)) return -1;
            const double rel = sqrt(c->hscal[S_RR]) / r0n;
            if (!(rel == rel)) { c->err = "Poisson FP32 CG produced NaN"; return -1; }
            if (rel <= tol_rel) break;
        }
        if (tb_vcycle32(c)) return -1;
        vcy++;
        if (rz(S_TMP0)) return -1;
        hipLaunchKernelGGL(k_cg32_p, dim3(PB), dim3(PT), 0, c->s, n4, (const float4 *)z, (float4 *)p,
                           (const double *)c->scal);
        hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(64), 0, c->s, c->scal, S_RZ, S_TMP0);
        TB_LAUNCH_CHECK(c);
    }
    return 0;
}

// Mixed precision: FP64 iterative refinement (true residual r = b - A x in
// FP64, rounded to FP32) around the FP32 inner CG, x += e in FP64.  Each inner
// solve asks for the residual reduction still missing (times 0.5), but never
// less than c->ir_floor (default 1e-5, which FP32 reaches reliably).  Stops at
// rtol, at maxit inner iterations, or when an outer pass reduces the FP64
// residual by less than half (stalled: the caller falls back to FP64 PCG).
extern "C" int tb_set_mixed_inner(tb_ctx *c, double reduction_floor, int maxit) {
    if (!c || !(reduction_floor > 0.0 && reduction_floor < 1.0) || maxit < 0) {
        if (c) c->err = "tb_set_mixed_inner: bad argument";
        return -1;
    }
    c->ir_floor = reduction_floor;
    c->ir_maxit = maxit;
    return 0;
}

// rr_known: x is already pinned (corner exactly 0) and lev[0].b32 holds its
// FP32 residual, of squared norm *rr_known; the first shift and residual are
// skipped.  Every refinement pass adds the correction and pins x in one
// kernel, so its residual (Neumann rows, with x[corner] = 0 the same vector
// as the pinned one) is the true residual of the result: no further pass
// follows the last one.  On return x is pinned, res is its pinned relative
// residual and the last residual pass ran on x.
static int mixed_ir(tb_ctx *c, const PoisCtx &P, double *x, const double *b, double bn, double rtol, int maxit,
                    int &its, int &vcy, double &res, bool &stalled, const double *rr_known) {
    double *r = nullptr;   // only the FP32 copy of the residual and its norms are needed
    Level &L0 = c->lev[0];
    double rr, bb;
    stalled = false;
    if (tb_mg_build(c)) return -1;
    const KTiles kt0 = make_ktiles(L0.nx, L0.ny, L0.nz);
    if (kt0.per_xcd * NXCD > RED_MAX_BLOCKS) { c->err = "grid too large for the reduction buffer"; return -1; }
    if (rr_known) {
        rr = *rr_known;
    } else {
        if (shift_corner(c, P, x)) return -1;
        if (residual(c, P, true, x, b, r, L0.b32, rr, bb)) return -1;
    }
    res = sqrt(rr) / bn;
    if (c->ir_maxit == 0 && res > rtol) {
        // no FP32 work allowed: the first pass would leave x as it is
        stalled = true;
        return 0;
    }
    while (res > rtol && its < maxit) {
        const double prev = res;
        double want = 0.5 * rtol / res;
        if (want < c->ir_floor) want = c->ir_floor;
        int room = maxit - its;
        // the first pass skips the tests of all but the last two iterations
        // the previous solve's first pass needed
        const int skip = (its == 0 && c->ir_first_its > 2) ? c->ir_first_its - 2 : 0;
        const int its_before = its;
        if (cg32_inner(c, kt0, sqrt(rr), want, room < c->ir_maxit ? room : c->ir_maxit, skip, its, vcy)) return -1;
        if (its_before == 0) c->ir_first_its = its;
        hipLaunchKernelGGL(k_corner_after_add, dim3(1), dim3(64), 0, c->s, (const double *)x,
                           (const float *)c->cg32[2], P.corner, c->scal);
        hipLaunchKernelGGL(k_add32_pin, tiles_grid(P.cells), tiles_block(), 0, c->s, c->d, x,
                           (const float *)c->cg32[2], (const double *)c->scal, P.cells);
        TB_LAUNCH_CHECK(c);
        if (residual(c, P, true, x, b, r, L0.b32, rr, bb)) return -1;
        res = sqrt(rr) / bn;
        if (!(res == res)) { c->err = "Poisson IR produced NaN"; return -1; }
        if (res > 0.5 * prev) { stalled = true; break; }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Initial guess of the learned mode: the A-norm-optimal combination
//   x0 = sum_k a_k phi_{n-k} (k = 1..H) + a_d d
// of the last H solutions and the network's correction d (when a network is
// attached and its gate is open), i.e. the Galerkin solution of A x = b on
// that subspace.  The Galerkin matrix of the stored solutions is kept from
// solve to solve (ctx->gram: A phi_{n-k} = b_{n-k} holds to the solve
// tolerance, so phi_i . A phi_j is phi_i . b_j, formed when the later of the
// two was solved), so one pass forms the products phi_k . b with the new
// right-hand side and, with the network, the products with d and A d.
// Returns the true relative residual of x0 (pinned operator) in resg and the
// number of directions used in ndir (0: none usable; the caller then restores
// phi_{n-1}).

// The directions are the stored solutions rounded to FP32: the guess only
// has to be good to the ~1e-3 relative residual it starts the solve from,
// and the dot products and the combination, which stream every direction,
// move half the bytes.  Products and sums stay in FP64; every direction is
// pinned as its FP64 original was (zeros round to zeros).
struct Dirs {
    const float *g[GUESS_HIST];     // stored solutions, newest first
};

// phi_hist[0] = x and hist32[0] = (float) x in one pass
__global__ void __launch_bounds__(PT) k_hist_push(long n, const double *__restrict__ x, double *__restrict__ h64,
                                                  float *__restrict__ h32) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = x[i];
        h64[i] = v;
        h32[i] = (float)v;
    }
}

// per block: the H sums phi_k . b, then with the network's direction d and
// its image q = A d the H sums phi_k . q, d . b and d . q
template <int H, bool NET>
__global__ void __launch_bounds__(PT) k_guess_dots(long n, Dirs D, const double *__restrict__ b,
                                                   const double *__restrict__ d, const double *__restrict__ q,
                                                   double *part) {
    constexpr int NS = NET ? 2 * H + 2 : H;
    __shared__ double sh[NS * 16];
    double a[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) a[k] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double bi = b[i];
        double qi = 0.0;
        if (NET) {
            const double di = d[i];
            qi = q[i];
            a[2 * H] += di * bi;
            a[2 * H + 1] += di * qi;
        }
#pragma unroll
        for (int u = 0; u < H; u++) {
            const double g = (double)D.g[u][i];
            a[u] += g * bi;
            if (NET) a[H + u] += g * qi;
        }
    }
    block_sum<NS>(a, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NS; k++) part[(long)blockIdx.x * NS + k] = a[k];
}

// out[k] = sum over blocks of part[blk * ncol + k]: one workgroup per
// column, each thread a fixed stride of blocks, then a fixed-order tree in
// LDS, so the sums do not depend on scheduling
constexpr int SUM_COLS_THREADS = 256;
__global__ void __launch_bounds__(SUM_COLS_THREADS) k_sum_cols(const double *part, int nblk, int ncol,
                                                                double *out) {
    __shared__ double sh[SUM_COLS_THREADS];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int bI = threadIdx.x; bI < nblk; bI += SUM_COLS_THREADS) s += part[(long)bI * ncol + k];
    sh[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int w = SUM_COLS_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
}

struct Coefs {
    double a[GUESS_DIRS];
};

// x = sum_{u < M} a_u g_u (+ a_M d when d is given); M is a template
// argument so that all loads of a point are in flight together
template <int M>
__global__ void __launch_bounds__(PT) k_combine(long n, double *__restrict__ x, Dirs D,
                                                const double *__restrict__ d, Coefs C) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double g[M];
#pragma unroll
        for (int u = 0; u < M; u++) g[u] = (double)D.g[u][i];
        const double dv = d ? d[i] : 0.0;
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < M; u++) s += C.a[u] * g[u];
        if (d) s += C.a[M] * dv;
        x[i] = s;
    }
}

static void launch_combine(tb_ctx *c, long n, double *x, const Dirs &D, int m, const double *d, const Coefs &C) {
    switch (m) {
    case 1: hipLaunchKernelGGL(k_combine<1>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 2: hipLaunchKernelGGL(k_combine<2>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 3: hipLaunchKernelGGL(k_combine<3>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 4: hipLaunchKernelGGL(k_combine<4>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 5: hipLaunchKernelGGL(k_combine<5>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 6: hipLaunchKernelGGL(k_combine<6>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 7: hipLaunchKernelGGL(k_combine<7>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 8: hipLaunchKernelGGL(k_combine<8>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    case 9: hipLaunchKernelGGL(k_combine<9>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    default: hipLaunchKernelGGL(k_combine<10>, dim3(PB), dim3(PT), 0, c->s, n, x, D, d, C); break;
    }
}

// solve the m x m system M a = r by Gaussian elimination with partial
// pivoting; false when a pivot falls below 1e-12 of the largest entry
static bool small_solve(int m, const double M[GUESS_DIRS][GUESS_DIRS], const double *r, double *a) {
    double A[GUESS_DIRS][GUESS_DIRS + 1];
    double mx = 0.0;
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < m; j++) { A[i][j] = M[i][j]; mx = std::max(mx, std::fabs(M[i][j])); }
        A[i][m] = r[i];
    }
    if (!(mx > 0.0)) return false;
    for (int col = 0; col < m; col++) {
        int piv = col;
        for (int i = col + 1; i < m; i++)
            if (std::fabs(A[i][col]) > std::fabs(A[piv][col])) piv = i;
        if (!(std::fabs(A[piv][col]) > 1e-12 * mx)) return false;
        for (int j = 0; j <= m; j++) std::swap(A[col][j], A[piv][j]);
        for (int i = 0; i < m; i++) {
            if (i == col) continue;
            const double f = A[i][col] / A[col][col];
            for (int j = col; j <= m; j++) A[i][j] -= f * A[col][j];
        }
    }
    for (int i = 0; i < m; i++) a[i] = A[i][m] / A[i][i];
    return true;
}

template <int H>
static void launch_dots(tb_ctx *c, long n, const Dirs &D, const double *b, const double *d, const double *q) {
    if (d) hipLaunchKernelGGL((k_guess_dots<H, true>), dim3(PB), dim3(PT), 0, c->s, n, D, b, d, q, c->red);
    else hipLaunchKernelGGL((k_guess_dots<H, false>), dim3(PB), dim3(PT), 0, c->s, n, D, b, d, q, c->red);
}

// On return with ndir > 0, x is pinned and lev[0].b32 holds the FP32 copy of
// its residual, whose squared norm is rrg (and that of b bbg); with ndir = 0
// x is untouched.  c->guess_r / guess_nr receive the products phi_k . b for
// the Galerkin matrix of the next solve.
static int learned_guess(tb_ctx *c, const PoisCtx &P, double *x, const double *b, double &rrg, double &bbg,
                         int &ndir) {
    const long n = c->d.ntot;
    double *dd = c->kv[5], *q = c->kv[6];
    ndir = 0;
    c->net_used = false;
    int H = c->guess_hist;
    if (H > c->n_hist) H = c->n_hist;
    if (H > c->gram_n) H = c->gram_n;
    if (H < 1) return 0;
    Dirs D;
    for (int k = 0; k < H; k++) D.g[k] = c->hist32[k];
    // the network reads the previous two solutions
    const bool use_d = c->net && c->net_open && c->n_hist >= 2;
    if (use_d) {
        TB_CHECK(c, hipMemsetAsync(dd, 0, sizeof(double) * n, c->s));
        TB_CHECK(c, hipEventRecord(c->net_ev[0], c->s));
        if (tb_net_predict(c, dd, 1)) return -1;
        TB_CHECK(c, hipEventRecord(c->net_ev[1], c->s));
        hipLaunchKernelGGL(k_pois_km<0>, ktiles_grid(P.kt), tiles_block(), 0, c->s, P.Pn, (const double *)dd,
                           (const double *)nullptr, q, (float *)nullptr, c->red, P.kt);
        TB_LAUNCH_CHECK(c);
    }
    const double *dp = use_d ? dd : nullptr, *qp = use_d ? q : nullptr;
    if (c->prof_on) tb_prof_begin(c, TB_K_GUESS);
    switch (H) {
    case 1: launch_dots<1>(c, n, D, b, dp, qp); break;
    case 2: launch_dots<2>(c, n, D, b, dp, qp); break;
    case 3: launch_dots<3>(c, n, D, b, dp, qp); break;
    case 4: launch_dots<4>(c, n, D, b, dp, qp); break;
    case 5: launch_dots<5>(c, n, D, b, dp, qp); break;
    case 6: launch_dots<6>(c, n, D, b, dp, qp); break;
    case 7: launch_dots<7>(c, n, D, b, dp, qp); break;
    case 8: launch_dots<8>(c, n, D, b, dp, qp); break;
    case 9: launch_dots<9>(c, n, D, b, dp, qp); break;
    default: launch_dots<10>(c, n, D, b, dp, qp); break;
    }
    static_assert(GUESS_HIST == 10, "one dot-product instantiation per history length");
    TB_LAUNCH_CHECK(c);
    const int ns = use_d ? 2 * H + 2 : H;
    hipLaunchKernelGGL(k_sum_cols, dim3(ns), dim3(SUM_COLS_THREADS), 0, c->s, (const double *)c->red, PB, ns,
                       c->scal + S_SUB);
    if (c->prof_on) tb_prof_end(c, TB_K_GUESS);
    TB_LAUNCH_CHECK(c);
    if (read_s(c, S_SUB, ns)) return -1;
    const double *h = c->hscal + S_SUB;
    for (int k = 0; k < H; k++) c->guess_r[k] = h[k];
    c->guess_nr = H;
    const int m = use_d ? H + 1 : H;
    double M[GUESS_DIRS][GUESS_DIRS], r[GUESS_DIRS], a[GUESS_DIRS];
    for (int i = 0; i < H; i++) {
        r[i] = h[i];
        for (int j = 0; j < H; j++) M[i][j] = c->gram[i][j];
    }
    if (use_d) {
        for (int k = 0; k < H; k++) M[k][H] = M[H][k] = h[H + k];
        r[H] = h[2 * H];
        M[H][H] = h[2 * H + 1];
    }
    // the largest leading block that is well conditioned (the network's
    // direction comes last and is the first to go)
    for (int k = m; k >= 1; k--) {
        if (small_solve(k, M, r, a)) { ndir = k; break; }
    }
    if (ndir == 0) return 0;
    const bool with_d = use_d && ndir == m;
    if (with_d && m > 1) {
        // energies J = -a.r/2 of the Galerkin guesses with and without d
        double a2[GUESS_DIRS];
        if (small_solve(m - 1, M, r, a2)) {
            double j2 = 0.0, j3 = 0.0;
            for (int i = 0; i < m - 1; i++) j2 += a2[i] * r[i];
            for (int i = 0; i < m; i++) j3 += a[i] * r[i];
            c->net_J2 = -0.5 * j2;
            c->net_J3 = -0.5 * j3;
            c->net_used = true;
        }
    }
    Coefs C;
    for (int k = 0; k < GUESS_DIRS; k++) C.a[k] = k < ndir ? a[k] : 0.0;
    TB_PROF(c, TB_K_GUESS, launch_combine(c, n, x, D, with_d ? H : ndir, with_d ? (const double *)dd : nullptr, C));
    TB_LAUNCH_CHECK(c);
    // Every stored solution is pinned (each solve leaves its x with corner
    // exactly 0 and half-cell planes 0), so a combination of them is too and
    // the shift would subtract exactly 0; only the network's direction needs
    // it.
    if (with_d && shift_corner(c, P, x)) return -1;
    // With x[corner] = 0 the pinned and the Neumann residual are the same
    // vector (the rows differ only in the corner's own coefficient): one
    // pass gives the guess's residual norm, the norm of b and the FP32
    // residual the mixed solve starts from.
    if (residual(c, P, true, x, b, nullptr, c->lev[0].b32, rrg, bbg)) return -1;
    return 0;
}

// The history rotates by one solution (newest first): the Galerkin matrix
// gains the row of the new solution, phi_0 . A phi_0 = phi_0 . b (xb) and
// phi_k . A phi_0 = phi_k . b (the guess's products, c->guess_r), and keeps
// the block of the older ones.
static void gram_push(tb_ctx *c, double xb) {
    double G[GUESS_HIST][GUESS_HIST];
    const int keep = std::min(c->gram_n, GUESS_HIST - 1);
    const int nr = std::min(c->guess_nr, keep);
    for (int i = 0; i < keep; i++)
        for (int j = 0; j < keep; j++) G[i + 1][j + 1] = c->gram[i][j];
    G[0][0] = xb;
    for (int k = 0; k < nr; k++) G[0][k + 1] = G[k + 1][0] = c->guess_r[k];
    const int n = 1 + nr;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) c->gram[i][j] = G[i][j];
    c->gram_n = n;
    c->guess_nr = 0;
}

extern "C" int tb_set_guess_history(tb_ctx *c, int h) {
    if (!c || h < 1 || h > GUESS_HIST) {
        if (c) c->err = "tb_set_guess_history: h must be 1..10";
        return -1;
    }
    c->guess_hist = h;
    return 0;
}

// After a solve that used the network's direction: its A-norm error
// reduction is sqrt(E2/E3) with E = 2 (J - J*), J* = -b.x/2 at the solution;
// divided by the solve's own per-iteration reduction that gives the
// iterations it saved.  After NET_TRIALS such solves the network stays on
// only if the saved iterations are worth more time than its inference.
constexpr int NET_TRIALS = 4;

static int judge_network(tb_ctx *c, const double *x, const double *b, double res0, double res, int its,
                         double solve_ms) {
    if (!c->net_used || c->net_trials >= NET_TRIALS) return 0;
    if (tb_dot(c, b, x, S_TMP0)) return -1;
    if (read_s(c, S_TMP0, 1)) return -1;
    float net_ms = 0.f;
    TB_CHECK(c, hipEventElapsedTime(&net_ms, c->net_ev[0], c->net_ev[1]));
    const double Js = -0.5 * c->hscal[S_TMP0];
    const double E2 = std::max(2.0 * (c->net_J2 - Js), 1e-300), E3 = std::max(2.0 * (c->net_J3 - Js), 1e-300);
    const double orders = 0.5 * std::log10(E2 / E3);
    const double per_it = (its > 0 && res > 0.0 && res0 > res) ? std::log10(res0 / res) / its : 1.0;
    c->net_saved += std::max(orders, 0.0) / per_it;
    c->net_cost_ms += net_ms;
    c->net_iter_ms += its > 0 ? (solve_ms - net_ms) / its : 0.0;
    c->net_trials++;
    if (c->net_trials == NET_TRIALS && c->net_saved * (c->net_iter_ms / NET_TRIALS) < c->net_cost_ms)
        c->net_open = false;
    return 0;
}

// state of that judgement: open (1) or closed (0), trials so far, mean
// iterations saved per solve and mean inference time per call
extern "C" int tb_net_gate(tb_ctx *c, int *open, int *trials, double *saved_per_solve, double *net_ms_per_call) {
    if (!c) return -1;
    const int t = c->net_trials;
    if (open) *open = c->net_open ? 1 : 0;
    if (trials) *trials = t;
    if (saved_per_solve) *saved_per_solve = t ? c->net_saved / t : 0.0;
    if (net_ms_per_call) *net_ms_per_call = t ? c->net_cost_ms / t : 0.0;
    return 0;
}

// One Poisson solve.  guess_f >= 0: the caller's initial guess in that
// field; output: the solve of Pressure_compute_real_pressure, which leaves
// the solution history, the Galerkin matrix, the snapshot store and the
// network's judgement as they are.
//
// A guess (the caller's, or the learned mode's Galerkin combination) is
// measured against the solve's plain start, the field as the previous solve
// left it (the reference's initial guess, Solver.c:256): it is kept when its
// residual is not larger; otherwise the previous field is restored and FP64
// GMG-PCG, the plain path, solves from it (fallback bit 2).  A mixed solve
// that stagnates hands over to FP64 GMG-PCG as well (fallback bit 1).
static int poisson_solve(tb_ctx *c, tb_solve_stats *st, int guess_f, bool output) {
    if (!c) return -1;
    if (tb_mg_build(c)) return -1;
    PoisCtx P = pois_ctx(c);
    if (P.nblk > RED_MAX_BLOCKS) { c->err = "grid too large for the reduction buffer"; return -1; }
    double *x = c->f[TB_F_PHI];
    const double *b = c->f[TB_F_PRHS];
    double rtol = c->cfg.pois_rtol;
    int maxit = c->cfg.pois_maxit;
    TB_CHECK(c, hipEventRecord(c->ev0, c->s));
    double rr = 0.0, bb = 0.0;
    const int mode = c->cfg.pois_mode;
    tb_solve_stats loc;
    memset(&loc, 0, sizeof(loc));
    int its = 0, vcy = 0;
    double res = 0.0;
    bool plain = mode == TB_POIS_FP64_PCG;
    c->net_used = false;
    const bool learned = !output && guess_f < 0 && mode == TB_POIS_LEARNED_IR && c->n_hist >= 1 && c->gram_n >= 1;
    // guess_resid: x holds an accepted guess and rr its residual, whose FP32
    // copy (mixed) or FP64 vector kv[0] (plain) the method starts from
    bool guess_resid = false;
    if (learned || guess_f >= 0) {
        // the plain start: the previous solution, which the learned mode
        // keeps as phi_hist[0]; any other field is saved (pinned) in kv[7]
        const double *prev = c->phi_is_hist0 ? c->phi_hist[0] : c->kv[7];
        if (!c->phi_is_hist0) {
            TB_CHECK(c, hipMemcpyAsync(c->kv[7], x, sizeof(double) * c->d.ntot, hipMemcpyDeviceToDevice, c->s));
            if (!c->phi_pinned && shift_corner(c, P, c->kv[7])) return -1;
        }
        double rr_g = 0.0, bb_g = 0.0;
        int ndir = 0;
        if (learned) {
            if (learned_guess(c, P, x, b, rr_g, bb_g, ndir)) return -1;
        } else {
            TB_CHECK(c, hipMemcpyAsync(x, c->f[guess_f], sizeof(double) * c->d.ntot, hipMemcpyDeviceToDevice, c->s));
            if (shift_corner(c, P, x)) return -1;
            if (residual(c, P, !plain, x, b, plain ? c->kv[0] : nullptr, plain ? nullptr : c->lev[0].b32, rr_g, bb_g))
                return -1;
            ndir = 1;
        }
        if (ndir > 0 && bb_g > 0.0) {
            // residual norm of the plain start (nothing written)
            double rr_p = 0.0, bb_p = 0.0;
            if (residual(c, P, true, prev, b, nullptr, nullptr, rr_p, bb_p)) return -1;
            if (rr_g <= rr_p) {
                guess_resid = true;
                rr = rr_g;
                bb = bb_g;
            } else {
                loc.fallback = 2;
                plain = true;
            }
        }
        if (!guess_resid && ndir > 0) {
            // back to the previous solution (pinned)
            TB_CHECK(c, hipMemcpyAsync(x, prev, sizeof(double) * c->d.ntot, hipMemcpyDeviceToDevice, c->s));
            c->phi_pinned = true;
        }
    }
    double rr_top = 0.0;
    if (!guess_resid) {
        // the field left by the previous solve is pinned already (the shift
        // would subtract exactly 0)
        if (!c->phi_pinned && shift_corner(c, P, x)) return -1;
        // this residual is also the first one of the method that follows:
        // FP64 PCG takes r (kv[0]), the mixed solve its FP32 copy (x is
        // pinned, so the Neumann and the pinned residual are the same vector)
        double *r_top = plain ? c->kv[0] : nullptr;
        float *r32_top = plain ? nullptr : c->lev[0].b32;
        if (residual(c, P, !plain, x, b, r_top, r32_top, rr, bb)) return -1;
    }
    rr_top = rr;
    c->phi_pinned = false;
    c->phi_is_hist0 = false;
    double bn = sqrt(bb);
    if (bn == 0.0) {
        TB_CHECK(c, hipMemsetAsync(x, 0, sizeof(double) * c->d.ntot, c->s));
        loc.converged = 1;
    } else {
        loc.res0 = sqrt(rr) / bn;
        if (plain) {
            if (pcg64(c, P, x, b, bn, rtol, maxit, its, vcy, res, &rr_top)) return -1;
        } else {
            bool stalled = false;
            if (mixed_ir(c, P, x, b, bn, rtol, maxit, its, vcy, res, stalled, &rr_top)) return -1;
            if (res > rtol) {
                loc.fallback |= 1;   // plain FP64 GMG-PCG takes over
                if (pcg64(c, P, x, b, bn, rtol, maxit, its, vcy, res, nullptr)) return -1;
            }
        }
        loc.converged = res <= rtol;
        if (c->net_used) {
            TB_CHECK(c, hipEventRecord(c->ev1, c->s));
            if (tb_sync(c)) return -1;
            if (judge_network(c, x, b, loc.res0, res, its, tb_elapsed_ms(c))) return -1;
        }
    }
    // every path above leaves x pinned (by its last shift or correction
    // pass, or all zero)
    c->phi_pinned = true;
    // history of solutions for the initial guess and the network's inputs,
    // newest first: the pointers rotate, one copy per solve; kept only where
    // something reads it, and not for output solves
    if (output) {
        // the history and its Galerkin matrix stay as they are
    } else if (c->cfg.pois_mode == TB_POIS_LEARNED_IR || c->net || c->net_record) {
        double *oldest = c->phi_hist[GUESS_HIST - 1];
        float *oldest32 = c->hist32[GUESS_HIST - 1];
        for (int k = GUESS_HIST - 1; k > 0; k--) {
            c->phi_hist[k] = c->phi_hist[k - 1];
            c->hist32[k] = c->hist32[k - 1];
        }
        c->phi_hist[0] = oldest;
        c->hist32[0] = oldest32;
        hipLaunchKernelGGL(k_hist_push, dim3(PB), dim3(PT), 0, c->s, c->d.ntot, (const double *)x, c->phi_hist[0],
                           c->hist32[0]);
        TB_LAUNCH_CHECK(c);
        if (c->n_hist < GUESS_HIST) c->n_hist++;
        c->phi_is_hist0 = true;
        // the last residual pass of every solve above ran on the final x
        // (x = 0 when b = 0) and left x . b in S_RES + 2
        if (c->cfg.pois_mode == TB_POIS_LEARNED_IR) gram_push(c, bn == 0.0 ? 0.0 : c->hscal[S_RES + 2]);
        else c->gram_n = c->guess_nr = 0;
    } else {
        c->n_hist = c->gram_n = c->guess_nr = 0;
    }
    if (!output && c->net_record && c->n_hist >= 3 && tb_net_record_one(c)) return -1;
    TB_CHECK(c, hipEventRecord(c->ev1, c->s));
    if (tb_sync(c)) return -1;
    loc.iterations = its;
    loc.vcycles = vcy;
    loc.res = (bn == 0.0) ? 0.0 : res;
    loc.ms = tb_elapsed_ms(c);
    if (st) *st = loc;
    return loc.converged ? 0 : 1;
}

extern "C" int tb_poisson_solve(tb_ctx *c, tb_solve_stats *st) { return poisson_solve(c, st, -1, false); }

extern "C" int tb_poisson_solve_guess(tb_ctx *c, int f_guess, tb_solve_stats *st) {
    if (!c || f_guess < 0 || f_guess >= TB_F_COUNT || f_guess == TB_F_PHI || f_guess == TB_F_PRHS) {
        if (c) c->err = "tb_poisson_solve_guess: bad field";
        return -1;
    }
    return poisson_solve(c, st, f_guess, false);
}

extern "C" int tb_real_pressure(tb_ctx *c, tb_solve_stats *st) {
    if (tb_real_pressure_rhs(c)) return -1;
    return poisson_solve(c, st, -1, true);
}

// ---------------------------------------------------------------------------
// Pressure_project_velocity (Pressure.c:280-439)
__global__ void __launch_bounds__(256) k_project(Dims d, GridDev g, double *__restrict__ u, double *__restrict__ v,
                                                 double *__restrict__ w, double *__restrict__ dpx,
                                                 double *__restrict__ dpy, double *__restrict__ dpz,
                                                 const double *__restrict__ phi, const double *__restrict__ b,
                                                 const uint8_t *__restrict__ su, const uint8_t *__restrict__ sv,
                                                 const uint8_t *__restrict__ sw, double dt, double coef,
                                                 double a_dEta, double a_dXi, double a_dPhi, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    double ph = phi[id], bc = b[id];
    if (su[id] != TB_BOUNDARY) {
        double metric_ex = g.mxu[i];
        double dd = (ph - phi[id - 1]) * a_dEta * metric_ex;
        u[id] -= dt * dd;
        double lap = -(bc - b[id - 1]) * metric_ex * a_dEta;
        dpx[id] += dd - coef * lap;
    }
    if (sv[id] != TB_BOUNDARY) {
        double C_y = a_dXi * g.myv[j];
        double dd = (ph - phi[id - d.px]) * C_y;
        v[id] -= dt * dd;
        double lap = -(bc - b[id - d.px]) * C_y;
        dpy[id] += dd - coef * lap;
    }
    if (sw[id] != TB_BOUNDARY) {
        double C_z = g.mzw[k] * a_dPhi;
        double dd = (ph - phi[id - d.plane]) * C_z;
        w[id] -= dt * dd;
        double lap = -(bc - b[id - d.plane]) * C_z;
        dpz[id] += dd - coef * lap;
    }
}

extern "C" int tb_project(tb_ctx *c, double dt) {
    if (!c) return -1;
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    if (c->prof_on) tb_prof_begin(c, TB_K_PROJECT);
    hipLaunchKernelGGL(k_project, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->g, c->f[TB_F_U], c->f[TB_F_V],
                       c->f[TB_F_W], c->f[TB_F_DPDX], c->f[TB_F_DPDY], c->f[TB_F_DPDZ], c->f[TB_F_PHI],
                       c->f[TB_F_PRHS], c->st[0], c->st[1], c->st[2], dt, dt / c->cfg.re, 1.0 / c->dEta,
                       1.0 / c->dXi, 1.0 / c->dPhi, t);
    if (c->prof_on) tb_prof_end(c, TB_K_PROJECT);
    TB_LAUNCH_CHECK(c);
    return 0;
}

// Pressure_compute_velocity_divergence (Pressure.c:462-586): |div| per cell
// and the maximum with the last (k, j, i)-ordered cell on ties.  The same
// pass is the NaN / Inf scan of Debugger_check_nan (Debugger.c:286-324) over
// the non-BOUNDARY cells: a cell whose divergence or concentration is not
// finite ranks as +infinity, so the maximum reports it with its index.
// Each workgroup walks the tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// (one tile when the grid covers them all); div may be null.
__global__ void __launch_bounds__(256) k_div(Dims d, GridDev g, const double *__restrict__ u,
                                             const double *__restrict__ v, const double *__restrict__ w,
                                             const double *__restrict__ cc, const uint8_t *__restrict__ sc,
                                             double *__restrict__ div, double *pmax, long *pidx, double a_dEta,
                                             double a_dXi, double a_dPhi, Tiles t, const double *__restrict__ uin,
                                             const double *__restrict__ uout) {
    __shared__ double sv[256];
    __shared__ long si[256];
    int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double best = -1.0;
    long bidx = -1;
    const long nvb = t.per_xcd * NXCD;
    for (long vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        int i, j, k;
        if (!tile_point_at(t, vb, i, j, k)) continue;
        long id = gidx(d, i, j, k);
        double dc = 0.0;
        bool finite = true;
        if (sc[id] != TB_BOUNDARY) {
            double dudx = (u[id + 1] - u[id]) * a_dEta * g.mxc[i];
            // u on an open x face from its plane (Pressure.c:541-548)
            if (uin && i == 0) dudx = (u[id + 1] - uin[(long)k * d.NY + j]) * a_dEta * g.mxc[i];
            if (uout && i == d.NX - 2) dudx = (uout[(long)k * d.NY + j] - u[id]) * a_dEta * g.mxc[i];
            double dvdy = (v[id + d.px] - v[id]) * a_dXi * g.myc[j];
            double dwdz = (w[id + d.plane] - w[id]) * a_dPhi * g.mzc[k];
            dc = fabs(dudx + dvdy + dwdz);
            // false for NaN and +-Inf
            finite = dc <= DBL_MAX && fabs(cc[id]) <= DBL_MAX;
        }
        if (div) div[id] = dc;
        const double key = finite ? dc : INFINITY;
        const long idx = ((long)k * d.NY + j) * d.NX + i;
        if (key > best || (key == best && idx > bidx)) { best = key; bidx = idx; }
    }
    sv[tid] = best;
    si[tid] = bidx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            double a = sv[tid], b2 = sv[tid + s];
            long ia = si[tid], ib = si[tid + s];
            if (b2 > a || (b2 == a && ib > ia)) { sv[tid] = b2; si[tid] = ib; }
        }
        __syncthreads();
    }
    if (tid == 0) { pmax[blockIdx.x] = sv[0]; pidx[blockIdx.x] = si[0]; }
}

__global__ void k_div_final(const double *pmax, const long *pidx, long n, double *out, long *oidx) {
    __shared__ double sv[1024];
    __shared__ long si[1024];
    double best = -1.0;
    long bidx = -1;
    for (long b = threadIdx.x; b < n; b += blockDim.x) {
        double a = pmax[b];
        long ia = pidx[b];
        if (a > best || (a == best && ia > bidx)) { best = a; bidx = ia; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bidx;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double a = sv[threadIdx.x], b2 = sv[threadIdx.x + s];
            long ia = si[threadIdx.x], ib = si[threadIdx.x + s];
            if (b2 > a || (b2 == a && ib > ia)) { sv[threadIdx.x] = b2; si[threadIdx.x] = ib; }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sv[0]; oidx[0] = si[0]; }
}

extern "C" int tb_div_max(tb_ctx *c, double *div_max, int32_t *ijk) {
    if (!c) return -1;
    Tiles t = make_tiles(c->d.NX, c->d.NY, c->d.NZ);
    long nb = t.per_xcd * NXCD;
    if (nb > RED_MAX_BLOCKS) { c->err = "grid too large for the reduction buffer"; return -1; }
    // per-block maxima and their indices in the first two partial-sum slots;
    // the result in S_TMP0 and, as the bits of a long, S_TMP1
    double *pm = c->red;
    long *pi = reinterpret_cast<long *>(c->red + RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_div, tiles_grid(t), tiles_block(), 0, c->s, c->d, c->g, c->f[TB_F_U], c->f[TB_F_V],
                       c->f[TB_F_W], c->f[TB_F_C], c->st[3], c->f[TB_F_DIV], pm, pi, 1.0 / c->dEta, 1.0 / c->dXi,
                       1.0 / c->dPhi, t, tb_in_plane(c, TB_Q_U), tb_out_plane(c, TB_Q_U));
    hipLaunchKernelGGL(k_div_final, dim3(1), dim3(1024), 0, c->s, pm, pi, nb, c->scal + S_TMP0,
                       reinterpret_cast<long *>(c->scal + S_TMP1));
    TB_LAUNCH_CHECK(c);
    if (tb_read_scalars(c, S_TMP0, 2)) return -1;
    const double h = c->hscal[S_TMP0];
    long hi;
    memcpy(&hi, &c->hscal[S_TMP1], sizeof(long));
    const int32_t ci = (int32_t)(hi % c->d.NX), cj = (int32_t)((hi / c->d.NX) % c->d.NY),
                  ck = (int32_t)(hi / ((long)c->d.NX * c->d.NY));
    if (div_max) *div_max = h;
    if (ijk) {
        ijk[0] = ci;
        ijk[1] = cj;
        ijk[2] = ck;
    }
    if (!(h <= DBL_MAX)) {
        char buf[160];
        snprintf(buf, sizeof(buf), "NaN or Inf in u, v, w or c at cell (i,j,k)=(%d,%d,%d)", ci, cj, ck);
        c->err = buf;
        return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// tb_div_max_noalloc: the same kernels on caller-owned device buffers, with
// the per-workgroup partials in a file-scope device array instead of memory
// the library allocates.  A fixed grid of NOALLOC_BLOCKS workgroups walks the
// tiles, so the partials take 2 * NOALLOC_BLOCKS slots whatever the box.
// Each GPU holds its own copy of the array until the process ends; every
// call on a GPU overwrites the slots its workgroups own before the final
// kernel reads them (nothing needs clearing), and calls that overlap on one
// GPU race for them.
constexpr int NOALLOC_BLOCKS = 2048;
constexpr long NOALLOC_MAX_POINTS = 1L << 31;
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS];   // maxima, then indices (bits of a long)

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static double *noalloc_ptr[NOALLOC_MAX_DEVICES];   // scratch's address on each device

// address of scratch on the current device, looked up on its first call
static int noalloc_scratch(double **p) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return -1;
        noalloc_ptr[dev] = static_cast<double *>(q);
    }
    *p = noalloc_ptr[dev];
    return 0;
}

extern "C" int tb_div_max_noalloc(const tb_box *box, const double *mxc, const double *myc, const double *mzc,
                                  const double *u, const double *v, const double *w, const double *c,
                                  const uint8_t *flags_c, double *div, double *out_max, int64_t *out_index,
                                  void *stream) {
    if (!box || !mxc || !myc || !mzc || !u || !v || !w || !c || !flags_c || !out_max || !out_index)
        return TB_NOALLOC_BAD_ARGUMENT;
    if (box->nx < 1 || box->ny < 1 || box->nz < 1 || box->pitch < box->nx + 1 || !(box->deta > 0.0) ||
        !(box->dxi > 0.0) || !(box->dphi > 0.0))
        return TB_NOALLOC_BAD_ARGUMENT;
    Dims d;
    d.NX = box->nx + 1;
    d.NY = box->ny + 1;
    d.NZ = box->nz + 1;
    d.px = box->pitch;
    d.plane = (long)d.px * d.NY;
    d.ntot = d.plane * d.NZ;
    if (d.ntot > NOALLOC_MAX_POINTS) return TB_NOALLOC_TOO_LARGE;
    double *scr = nullptr;
    if (noalloc_scratch(&scr)) return TB_NOALLOC_NO_SCRATCH;
    GridDev g;
    memset(&g, 0, sizeof(g));
    g.mxc = mxc;
    g.myc = myc;
    g.mzc = mzc;
    const Tiles t = make_tiles(d.NX, d.NY, d.NZ);
    const long nvb = t.per_xcd * NXCD;
    const int nb = nvb < NOALLOC_BLOCKS ? (int)nvb : NOALLOC_BLOCKS;
    double *pm = scr;
    long *pi = reinterpret_cast<long *>(scr + NOALLOC_BLOCKS);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_div, dim3(nb), tiles_block(), 0, st, d, g, u, v, w, c, flags_c, div, pm, pi,
                       1.0 / box->deta, 1.0 / box->dxi, 1.0 / box->dphi, t, (const double *)nullptr,
                       (const double *)nullptr);
    hipLaunchKernelGGL(k_div_final, dim3(1), dim3(1024), 0, st, (const double *)pm, (const long *)pi, (long)nb,
                       out_max, reinterpret_cast<long *>(out_index));
    if (hipGetLastError() != hipSuccess) return TB_NOALLOC_LAUNCH;
    return 0;
}
