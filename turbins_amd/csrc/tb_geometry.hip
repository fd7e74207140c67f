// tb_geometry.hip -- signed-distance reinitialisation, node classification
// and immersed-boundary node tables, all on the device.  The arithmetic of
// every kernel follows the reference statement order and the library is
// built with -ffp-contract=off, so flags and IB coefficients are
// bit-identical to a sequential FP64 evaluation of the same formulas.
#include <cmath>
#include <cstring>

#include "tb_common.h"

void tb_free_ib(tb_ctx *c, int q);

__device__ inline double dmax(double a, double b) { return a < b ? b : a; }   // max() of definitions.h:50
__device__ inline double dmin(double a, double b) { return a > b ? b : a; }
__device__ inline int imax(int a, int b) { return a < b ? b : a; }
__device__ inline int imin(int a, int b) { return a > b ? b : a; }

// Surface_initialize (Surface.c:96-192): sdf guess y - interface(x, z)
__global__ void k_sdf_guess(Dims d, GridDev g, const double *__restrict__ h, double offset,
                            double *__restrict__ p0, double *__restrict__ p, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    double v = g.yc[j] - (h[(long)k * d.NX + i] + offset);
    p0[id] = v;
    p[id] = v;
}

// Levelset_compute_derivatives (Levelset.c:163-261)
__global__ void k_ls_deriv(Dims d, GridDev g, const double *__restrict__ p, double *__restrict__ Dx,
                           double *__restrict__ Dy, double *__restrict__ Dz, double a_dEta, double a_dXi,
                           double a_dPhi, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
    long id = gidx(d, i, j, k);
    if (i < NI - 1) Dx[id] = (p[id + 1] - p[id]) * g.mxu[i] * a_dEta;
    else Dx[id] = (p[id] - p[id - 1]) * g.mxu[i - 1] * a_dEta;
    if (j < NJ - 1) Dy[id] = (p[id + d.px] - p[id]) * g.myv[j] * a_dXi;
    else Dy[id] = (p[id] - p[id - d.px]) * g.myv[j - 1] * a_dXi;
    if (k < NK - 1) Dz[id] = (p[id + d.plane] - p[id]) * g.mzw[k] * a_dPhi;
    else Dz[id] = (p[id] - p[id - d.plane]) * g.mzw[k - 1] * a_dPhi;
}

// Levelset_do_Godunov (Levelset.c:267-494), Russo-Smereka interface fix
__global__ void k_ls_godunov(Dims d, GridDev g, const double *__restrict__ p0, double *__restrict__ p,
                             const double *__restrict__ Dx, const double *__restrict__ Dy,
                             const double *__restrict__ Dz, double dt, int scaled_guard, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;   // box is [0,NI) x [0,NJ) x [0,NK)
    int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
    long id = gidx(d, i, j, k);
    double dx = g.dxc[i];
    double pp0 = p0[id];
    double sign_0 = pp0 / sqrt(pp0 * pp0 + dx * dx);
    double Dxp = Dx[id], Dxm = Dx[gidx(d, imax(0, i - 1), j, k)];
    double Dyp = Dy[id], Dym = Dy[gidx(d, i, imax(0, j - 1), k)];
    double Dzp = Dz[id], Dzm = Dz[gidx(d, i, j, imax(0, k - 1))];
    double ap = dmax(Dxm, 0.0), am = dmin(Dxm, 0.0);
    double bp = dmax(Dxp, 0.0), bm = dmin(Dxp, 0.0);
    double cp = dmax(Dym, 0.0), cm = dmin(Dym, 0.0);
    double dp = dmax(Dyp, 0.0), dm = dmin(Dyp, 0.0);
    double ep = dmax(Dzm, 0.0), em = dmin(Dzm, 0.0);
    double fp = dmax(Dzp, 0.0), fm = dmin(Dzp, 0.0);
    double tx, ty, tz;
    if (sign_0 > 0.0) {
        tx = dmax(ap * ap, bm * bm); ty = dmax(cp * cp, dm * dm); tz = dmax(ep * ep, fm * fm);
    } else {
        tx = dmax(am * am, bp * bp); ty = dmax(cm * cm, dp * dp); tz = dmax(em * em, fp * fp);
    }
    double H = sign_0 * (sqrt(tx + ty + tz) - 1.0);
    int ix;
    ix = imin(i + 1, NI - 1); double p_e = p0[gidx(d, ix, j, k)]; double DeltaX = g.xc[ix];
    ix = imin(j + 1, NJ - 1); double p_n = p0[gidx(d, i, ix, k)]; double DeltaY = g.yc[ix];
    ix = imin(k + 1, NK - 1); double p_f = p0[gidx(d, i, j, ix)]; double DeltaZ = g.zc[ix];
    ix = imax(i - 1, 0); double p_w = p0[gidx(d, ix, j, k)]; DeltaX -= g.xc[ix];
    ix = imax(j - 1, 0); double p_s = p0[gidx(d, i, ix, k)]; DeltaY -= g.yc[ix];
    ix = imax(k - 1, 0); double p_b = p0[gidx(d, i, j, ix)]; DeltaZ -= g.zc[ix];
    double p_p = pp0;
    if (((p_p * p_w) < 0.0) || ((p_p * p_e) < 0.0) || ((p_p * p_n) < 0.0) || ((p_p * p_s) < 0.0) ||
        ((p_p * p_b) < 0.0) || ((p_p * p_f) < 0.0)) {
        const double tol = 0.01;
        // scaled guard: min(tol, 1.5 d_c) keeps interior central differences
        // on fine grids; equal to the absolute tol when every d_c >= tol/1.5
        double gx = tol, gy = tol, gz = tol;
        if (scaled_guard) {
            gx = dmin(tol, 1.5 * g.dxc[i]);
            gy = dmin(tol, 1.5 * g.dyc[j]);
            gz = dmin(tol, 1.5 * g.dzc[k]);
        }
        double Dphi_dx = (fabs(DeltaX) > gx) ? (p_e - p_w) / DeltaX : 0.0;
        double Dphi_dy = (fabs(DeltaY) > gy) ? (p_n - p_s) / DeltaY : 0.0;
        double Dphi_dz = (fabs(DeltaZ) > gz) ? (p_f - p_b) / DeltaZ : 0.0;
        double Dt = sqrt(Dphi_dx * Dphi_dx + Dphi_dy * Dphi_dy + Dphi_dz * Dphi_dz);
        if (Dt < tol) Dt = tol;
        double D = p_p / Dt;
        if (fabs(sign_0) < 1.0e-8) sign_0 = 0.0;
        else if (sign_0 > 0.0) sign_0 = 1.0;
        else sign_0 = -1.0;
        double ddx = g.dxc[i], ddy = g.dyc[j], ddz = g.dzc[k];
        double Ds = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        H = (sign_0 * fabs(p[id]) - D) / Ds;
    }
    p[id] += -H * dt;
}

// Levelset_is_converged, NORM2 (Levelset.c:499-576): partial sums of
// error^2 dx dy dz, then pold <- p (Levelset_store_old_data).
__global__ void k_ls_norm(Dims d, GridDev g, const double *__restrict__ p, double *__restrict__ pold,
                          double *part, Tiles t) {
    __shared__ double sh[16];
    double v[1] = {0.0};
    int i, j, k;
    if (tile_point(t, i, j, k)) {
        long id = gidx(d, i, j, k);
        double error = fabs(p[id] - pold[id]);
        v[0] = error * error * g.dxc[i] * g.dyc[j] * g.dzc[k];
        pold[id] = p[id];
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        long b = blockIdx.x;
        part[b] = v[0];
    }
}

__global__ void k_sum_partials(const double *part, long n, double *out) {
    __shared__ double sh[16];
    double v[1] = {0.0};
    for (long b = threadIdx.x; b < n; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) out[0] = v[0];
}

// Levelset_update_last_planes (Levelset.c:642-709), one plane per launch
__global__ void k_ls_last_x(Dims d, double *p) {
    long n = (long)d.NY * d.NZ;
    long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= n) return;
    int j = (int)(a % d.NY), k = (int)(a / d.NY);
    p[gidx(d, d.NX - 1, j, k)] = p[gidx(d, d.NX - 2, j, k)];
}
__global__ void k_ls_last_y(Dims d, double *p) {
    long n = (long)d.NX * d.NZ;
    long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= n) return;
    int i = (int)(a % d.NX), k = (int)(a / d.NX);
    p[gidx(d, i, d.NY - 1, k)] = p[gidx(d, i, d.NY - 2, k)];
}
__global__ void k_ls_last_z(Dims d, double *p) {
    long n = (long)d.NX * d.NY;
    long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= n) return;
    int i = (int)(a % d.NX), j = (int)(a / d.NX);
    p[gidx(d, i, j, d.NZ - 1)] = p[gidx(d, i, j, d.NZ - 2)];
}

// MyMath_interpolate_quantity, two points (MyMath.c:24-41)
__device__ inline double interp2(double q0, double x0, double q1, double x1, double xi) {
    double m = (q1 - q0) / (x1 - x0);
    if (fabs(x1 - x0) < 1.0e-8) return q0;
    return m * (xi - x0) + q0;
}

// Surface_find_q_sdf (Surface.c:216-334)
__global__ void k_q_sdf(Dims d, GridDev g, const double *__restrict__ cs, double *__restrict__ us,
                        double *__restrict__ vs, double *__restrict__ ws, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    double c0 = cs[id];
    double q1, x1;
    if (i > 0) { q1 = cs[id - 1]; x1 = g.xc[i - 1]; } else { q1 = c0; x1 = g.xc[i]; }
    us[id] = interp2(c0, g.xc[i], q1, x1, g.xu[i]);
    if (j > 0) { q1 = cs[id - d.px]; x1 = g.yc[j - 1]; } else { q1 = c0; x1 = g.yc[j]; }
    vs[id] = interp2(c0, g.yc[j], q1, x1, g.yv[j]);
    if (k > 0) { q1 = cs[id - d.plane]; x1 = g.zc[k - 1]; } else { q1 = c0; x1 = g.zc[k]; }
    ws[id] = interp2(c0, g.zc[k], q1, x1, g.zw[k]);
}

__device__ inline int isgn(double a) { return (a > 0.0) ? 1 : ((a < 0.0) ? -1 : 0); }

// Grid_tag_q_box_boundary_nodes (Grid.c:1617-1685) and
// Grid_tag_q_nodes_using_surface (Grid.c:1424-1612)
__global__ void k_classify(Dims d, int q, int has_geo, const double *__restrict__ sdf, uint8_t *__restrict__ st,
                           Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    int NX = d.NX, NY = d.NY, NZ = d.NZ;
    int NI = NX - 1, NJ = NY - 1, NK = NZ - 1;
    long id = gidx(d, i, j, k);
    bool bnd = (i == NX - 1) || (j == NY - 1) || (k == NZ - 1) || (q == TB_Q_U && i == 0) ||
               (q == TB_Q_V && j == 0) || (q == TB_Q_W && k == 0);
    if (bnd) { st[id] = TB_BOUNDARY; return; }
    if (!has_geo) { st[id] = TB_FLUID; return; }
    double s_p = sdf[id];
    if (s_p >= 0.0) { st[id] = TB_FLUID; return; }
    double s_e = sdf[gidx(d, imin(i + 1, NI - 1), j, k)];
    double s_w = sdf[gidx(d, imax(i - 1, 0), j, k)];
    double s_n = sdf[gidx(d, i, imin(j + 1, NJ - 1), k)];
    double s_s = sdf[gidx(d, i, imax(j - 1, 0), k)];
    double s_f = sdf[gidx(d, i, j, imin(k + 1, NK - 1))];
    double s_b = (double)isgn(sdf[gidx(d, i, j, imax(k - 1, 0))]);
    bool imm = ((s_p * s_e) < 0.0) || ((s_p * s_w) < 0.0) || ((s_p * s_n) < 0.0) || ((s_p * s_s) < 0.0) ||
               ((s_p * s_b) < 0.0) || ((s_p * s_f) < 0.0);
    st[id] = imm ? TB_IMMERSED : TB_SOLID;
}

// Grid_tag_q_buffer_solid_nodes (Grid.c:2024-2162), width 2, reading the
// pre-pass statuses (the ghosted local copy) and writing the result.
__global__ void k_buffer(Dims d, const uint8_t *__restrict__ pre, uint8_t *__restrict__ st, Tiles t) {
    int i, j, k;
    if (!tile_point(t, i, j, k)) return;
    long id = gidx(d, i, j, k);
    if (pre[id] != TB_SOLID) return;
    const int w = 2;
    for (int kz = -w; kz <= w; kz++)
        for (int ky = -w; ky <= w; ky++)
            for (int kx = -w; kx <= w; kx++) {
                int in = i + kx, jn = j + ky, kn = k + kz;
                if (in < 0 || in > d.NX - 1 || jn < 0 || jn > d.NY - 1 || kn < 0 || kn > d.NZ - 1) continue;
                if (pre[gidx(d, in, jn, kn)] == TB_IMMERSED) { st[id] = TB_BUFFER_SOLID; return; }
            }
}

// ---------------------------------------------------------------------------
// Immersed-boundary node setup (Immersed.c), one thread per IB node.

struct QCoords { const double *x, *y, *z; };

__device__ inline int q_status(const Dims &d, const uint8_t *st, int i, int j, int k) {
    if (i < 0 || i > d.NX - 1 || j < 0 || j > d.NY - 1 || k < 0 || k > d.NZ - 1) return TB_OUTSIDE;
    return st[gidx(d, i, j, k)];
}

__device__ inline double pdist(const double *a, const double *b) {
    double d2 = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    return sqrt(d2);
}

// MyMath_binary_search (MyMath.c:384-398)
__device__ int bsearch_idx(const double *a, double v, int lo, int hi) {
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (v >= a[mid]) {
            if (v < a[mid + 1]) return mid;
            lo = mid + 1;
        } else if (v < a[mid]) {
            hi = mid - 1;
        }
    }
    return -1;
}

// Grid_get_x/y/z_index (Grid.c:1877-1967); the z test is ">=".
__device__ int grid_index(const double *xq, double x, int N, bool zdir) {
    if (x < xq[0]) return -1;
    if (zdir) { if (x >= xq[N - 1]) return 10000000; }
    else if (x > xq[N - 1]) return 10000000;
    return bsearch_idx(xq, x, 0, N - 1);
}

// MyMath_GaussJordan_matrix_inv (MyMath.c:117-209) for n = 8
__device__ void gauss_jordan8(double A[8][8]) {
    int index_col[8], index_row[8], index_piv[8];
    int in_row = -1, in_col = -1;
    for (int j = 0; j < 8; j++) index_piv[j] = -1;
    for (int i = 0; i < 8; i++) {
        double mx = 0.0;
        for (int j = 0; j < 8; j++)
            if (index_piv[j] != 0)
                for (int k = 0; k < 8; k++)
                    if (index_piv[k] == -1 && fabs(A[j][k]) >= mx) { mx = fabs(A[j][k]); in_row = j; in_col = k; }
        ++(index_piv[in_col]);
        if (in_row != in_col)
            for (int h = 0; h < 8; h++) { double t = A[in_row][h]; A[in_row][h] = A[in_col][h]; A[in_col][h] = t; }
        index_row[i] = in_row;
        index_col[i] = in_col;
        double a_piv = 1.0 / A[in_col][in_col];
        A[in_col][in_col] = 1.0;
        for (int h = 0; h < 8; h++) A[in_col][h] *= a_piv;
        for (int hh = 0; hh < 8; hh++)
            if (hh != in_col) {
                double t = A[hh][in_col];
                A[hh][in_col] = 0.0;
                for (int h = 0; h < 8; h++) A[hh][h] -= A[in_col][h] * t;
            }
    }
    for (int h = 7; h >= 0; h--)
        if (index_row[h] != index_col[h])
            for (int k = 0; k < 8; k++) {
                double t = A[k][index_row[h]];
                A[k][index_row[h]] = A[k][index_col[h]];
                A[k][index_col[h]] = t;
            }
}

__global__ void k_ib_nodes(Dims d, GridDev g, int q, QCoords qc, const double *__restrict__ sdf,
                           const uint8_t *__restrict__ st, tb_ib_node *nodes, int n) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    tb_ib_node nd = nodes[a];   // i, j, k filled by the host compaction
    int i = nd.i, j = nd.j, k = nd.k;
    const double *xq = qc.x, *yq = qc.y, *zq = qc.z;
    nd.im[0] = xq[i]; nd.im[1] = yq[j]; nd.im[2] = zq[k];
    int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
    if (q == TB_Q_U) NI++;
    if (q == TB_Q_V) NJ++;
    if (q == TB_Q_W) NK++;
    // Surface_compute_normal (Surface.c:547-619)
    int iw = imax(i - 1, 0), ie = imin(i + 1, NI - 1);
    int is = imax(j - 1, 0), in = imin(j + 1, NJ - 1);
    int ib = imax(k - 1, 0), iff = imin(k + 1, NK - 1);
    double Delta_x = xq[ie] - xq[iw];
    double Delta_y = yq[in] - yq[is];
    double Delta_z = zq[iff] - zq[ib];
    double Dsx = (sdf[gidx(d, ie, j, k)] - sdf[gidx(d, iw, j, k)]) / Delta_x;
    double Dsy = (sdf[gidx(d, i, in, k)] - sdf[gidx(d, i, is, k)]) / Delta_y;
    double Dsz = (sdf[gidx(d, i, j, iff)] - sdf[gidx(d, i, j, ib)]) / Delta_z;
    double a_norm = 1.0 / sqrt(Dsx * Dsx + Dsy * Dsy + Dsz * Dsz);
    nd.n[0] = Dsx * a_norm; nd.n[1] = Dsy * a_norm; nd.n[2] = Dsz * a_norm;
    // Immersed_find_control_point + Surface_find_Dn_to_surface
    double D_sdf_dn = nd.n[0] * Dsx + nd.n[1] * Dsy + nd.n[2] * Dsz;
    if ((fabs(D_sdf_dn) < 1.0e-2) || isnan(D_sdf_dn)) D_sdf_dn = 1.0e-2;
    double Dn = sdf[gidx(d, i, j, k)] / D_sdf_dn;
    nd.control[0] = nd.im[0] - nd.n[0] * Dn;
    nd.control[1] = nd.im[1] - nd.n[1] * Dn;
    nd.control[2] = nd.im[2] - nd.n[2] * Dn;
    // Immersed_find_image_point (Immersed.c:1063-1149)
    double ds = pdist(nd.im, nd.control);
    double D = 2.0 * ds;
    nd.image[0] = nd.im[0] + nd.n[0] * D;
    nd.image[1] = nd.im[1] + nd.n[1] * D;
    nd.image[2] = nd.im[2] + nd.n[2] * D;
    // Immersed_find_box (Immersed.c:1210-1402)
    const int dx8[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    const int dy8[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    const int dz8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double x = nd.image[0], y = nd.image[1], z = nd.image[2];
    const double tol_x = dmax(0.005 * g.dxc[i], 1e-12);
    const double tol_y = dmax(0.005 * g.dyc[j], 1e-12);
    const double tol_z = dmax(0.005 * g.dzc[k], 1e-12);
    const double px[8] = {+tol_x, -tol_x, +tol_x, -tol_x, +tol_x, -tol_x, +tol_x, -tol_x};
    const double py[8] = {+tol_y, +tol_y, -tol_y, -tol_y, +tol_y, +tol_y, -tol_y, -tol_y};
    const double pz[8] = {+tol_z, +tol_z, +tol_z, +tol_z, -tol_z, -tol_z, -tol_z, -tol_z};
    int m = 0, found = 0, i_ig = 0, j_ig = 0, k_ig = 0;
    int Nx = d.NX - 1, Ny = d.NY - 1, Nz = d.NZ - 1;   // grid->NI, NJ, NK
    while (!found && m < 8) {
        i_ig = grid_index(xq, x + px[m], Nx, false);
        j_ig = grid_index(yq, y + py[m], Ny, false);
        k_ig = grid_index(zq, z + pz[m], Nz, true);
        int good = 1;
        for (int c8 = 0; c8 < 8; c8++) {
            int s = q_status(d, st, i_ig + dx8[c8], j_ig + dy8[c8], k_ig + dz8[c8]);
            if (s != TB_FLUID && s != TB_IMMERSED) { good = 0; break; }
        }
        found = good;
        if (found && !((x >= xq[i_ig]) && (x <= xq[i_ig + 1]) && (y >= yq[j_ig]) && (y <= yq[j_ig + 1]) &&
                       (z >= zq[k_ig]) && (z <= zq[k_ig + 1])))
            found = 0;
        m++;
    }
    double fp[8][3];
    for (int c8 = 0; c8 < 8; c8++) {
        nd.fi[c8] = nd.fj[c8] = nd.fk[c8] = 0;
        nd.coef[c8] = 0.0;
        fp[c8][0] = fp[c8][1] = fp[c8][2] = 0.0;
    }
    if (found) {
        nd.n_fluid = 8;
        for (int c8 = 0; c8 < 8; c8++) {
            int i_f = i_ig + dx8[c8], j_f = j_ig + dy8[c8], k_f = k_ig + dz8[c8];
            nd.fi[c8] = i_f; nd.fj[c8] = j_f; nd.fk[c8] = k_f;
            fp[c8][0] = xq[i_f]; fp[c8][1] = yq[j_f]; fp[c8][2] = zq[k_f];
        }
    } else {
        double D_min = 1e12;
        int g_found = -1;
        i_ig = grid_index(xq, x, Nx, false);
        j_ig = grid_index(yq, y, Ny, false);
        k_ig = grid_index(zq, z, Nz, true);
        i_ig = imax(i_ig, 0); j_ig = imax(j_ig, 0); k_ig = imax(k_ig, 0);
        i_ig = imin(i_ig, Nx - 1); j_ig = imin(j_ig, Ny - 1); k_ig = imin(k_ig, Nz - 1);
        double p3[3] = {x, y, z};
        for (int c8 = 0; c8 < 8; c8++) {
            int i_ = i_ig + dx8[c8], j_ = j_ig + dy8[c8], k_ = k_ig + dz8[c8];
            if (q_status(d, st, i_, j_, k_) == TB_FLUID) {
                double pp[3] = {xq[i_], yq[j_], zq[k_]};
                double dd = pdist(p3, pp);
                if (dd < D_min) { D_min = dd; g_found = c8; }
                found = 1;
            }
        }
        if (found) {
            nd.n_fluid = 1;
            int bi = i_ig + dx8[g_found], bj = j_ig + dy8[g_found], bk = k_ig + dz8[g_found];
            nd.fi[0] = bi; nd.fj[0] = bj; nd.fk[0] = bk;
            fp[0][0] = xq[bi]; fp[0][1] = yq[bj]; fp[0][2] = zq[bk];
        } else {
            nd.n_fluid = 0;
        }
    }
    // Immersed_compute_interpolation_coef (Immersed.c:122-374)
    double df = pdist(nd.image, nd.control);
    ds = pdist(nd.im, nd.control);
    double coef = ((fabs(ds) <= 1e-10) && (fabs(df) <= 1e-10)) ? 1.0 : ds / df;
    double wgt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (nd.n_fluid == 8) {
        double x7 = fp[7][0], y7 = fp[7][1], z7 = fp[7][2];
        double x0 = fp[0][0], y0 = fp[0][1], z0 = fp[0][2];
        double alpha = (x7 - nd.image[0]) / (x7 - x0);
        double beta = (y7 - nd.image[1]) / (y7 - y0);
        double gamma = (z7 - nd.image[2]) / (z7 - z0);
        wgt[0] = alpha * beta * gamma;
        wgt[1] = (1.0 - alpha) * beta * gamma;
        wgt[2] = alpha * (1.0 - beta) * gamma;
        wgt[3] = (1.0 - alpha) * (1.0 - beta) * gamma;
        wgt[4] = alpha * beta * (1.0 - gamma);
        wgt[5] = (1.0 - alpha) * beta * (1.0 - gamma);
        wgt[6] = alpha * (1.0 - beta) * (1.0 - gamma);
        wgt[7] = (1.0 - alpha) * (1.0 - beta) * (1.0 - gamma);
    } else if (nd.n_fluid == 1) {
        wgt[0] = 1.0;
    }
    if (q != TB_Q_C) {   // DIRICHLET for u, v, w
        nd.boundary_coef = 1.0 + coef;
        for (int c8 = 0; c8 < nd.n_fluid; c8++) nd.coef[c8] = wgt[c8] * coef * (+1.0);
    } else if (nd.n_fluid == 1) {   // NEUMANN, one node
        nd.boundary_coef = -(ds + df);
        nd.coef[0] = wgt[0] * (-1.0);
    } else {                          // NEUMANN, method 3
        double A[8][8];
        for (int r = 0; r < 8; r++) for (int cc = 0; cc < 8; cc++) A[r][cc] = 0.0;
        int self_included = 0, r = 0;
        for (int c8 = 0; c8 < nd.n_fluid; c8++) {
            if (nd.fi[c8] == i && nd.fj[c8] == j && nd.fk[c8] == k) {
                self_included = 1;
            } else {
                double xx = fp[c8][0] - nd.image[0], yy = fp[c8][1] - nd.image[1], zz = fp[c8][2] - nd.image[2];
                A[r][0] = xx * yy * zz; A[r][1] = xx * yy; A[r][2] = xx * zz; A[r][3] = yy * zz;
                A[r][4] = xx; A[r][5] = yy; A[r][6] = zz; A[r][7] = 1.0;
                r++;
            }
        }
        if (self_included) {
            double xc = nd.control[0] - nd.image[0], yc = nd.control[1] - nd.image[1], zc = nd.control[2] - nd.image[2];
            double nx = nd.n[0], ny = nd.n[1], nz = nd.n[2];
            A[r][0] = nx * yc * zc + ny * xc * zc + nz * xc * yc;
            A[r][1] = yc * nx + xc * ny;
            A[r][2] = zc * nx + xc * nz;
            A[r][3] = zc * ny + yc * nz;
            A[r][4] = nx; A[r][5] = ny; A[r][6] = nz; A[r][7] = 0.0;
        }
        if (nd.n_fluid == 8) gauss_jordan8(A);
        double rhs = self_included ? A[7][7] : 0.0;
        nd.boundary_coef = -(ds + df) + rhs;
        r = 0;
        for (int c8 = 0; c8 < nd.n_fluid; c8++) {
            if (nd.fi[c8] == i && nd.fj[c8] == j && nd.fk[c8] == k) nd.coef[c8] = 0.0;
            else { nd.coef[c8] = -A[7][r]; r++; }
        }
    }
    for (int c8 = 0; c8 < nd.n_fluid; c8++)
        if (isnan(nd.coef[c8]) || isinf(nd.coef[c8])) { nd.coef[c8] = 0.0; nd.boundary_coef = 0.0; }
    nodes[a] = nd;
}

// Matrix row of an IMMERSED node (Velocity.c:718-762, Conc.c:668-708):
// self coefficient 1 + sum of self weights, other weights with
// |coef| >= LSYS_DROP_TOLERANCE, all scaled by ap_solid; columns sorted.
__global__ void k_ib_rows(Dims d, const tb_ib_node *nodes, int n, double ap_solid, long *row, int *cnt, long *col,
                          double *val) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const tb_ib_node &nd = nodes[a];
    long cc[9];
    double vv[9];
    int m = 0;
    long self = gidx(d, nd.i, nd.j, nd.k);
    cc[m] = self; vv[m] = 1.0; m++;
    for (int g8 = 0; g8 < nd.n_fluid; g8++) {
        if (nd.fi[g8] == nd.i && nd.fj[g8] == nd.j && nd.fk[g8] == nd.k) vv[0] += nd.coef[g8];
        else if (fabs(nd.coef[g8]) >= 1e-8) { cc[m] = gidx(d, nd.fi[g8], nd.fj[g8], nd.fk[g8]); vv[m] = nd.coef[g8]; m++; }
    }
    for (int s = 0; s < m; s++) vv[s] *= ap_solid;
    for (int p = 1; p < m; p++) {
        long ck = cc[p]; double vk = vv[p]; int b = p - 1;
        while (b >= 0 && cc[b] > ck) { cc[b + 1] = cc[b]; vv[b + 1] = vv[b]; b--; }
        cc[b + 1] = ck; vv[b + 1] = vk;
    }
    row[a] = self;
    cnt[a] = m;
    for (int s = 0; s < 9; s++) {
        col[(long)a * 9 + s] = (s < m) ? cc[s] : self;
        val[(long)a * 9 + s] = (s < m) ? vv[s] : 0.0;
    }
}

// ---------------------------------------------------------------------------
static double ap_solid_for(tb_ctx *c, int q) {
    double kappa = (q == TB_Q_C) ? c->cfg.pe : c->cfg.re;
    double Cx = -1.0 / (kappa * c->dEta * c->dEta);
    double Cy = -1.0 / (kappa * c->dXi * c->dXi);
    double Cz = -1.0 / (kappa * c->dPhi * c->dPhi);
    return -Cx * c->hgrid[15][0] - Cy * c->hgrid[16][0] - Cz * c->hgrid[17][0];
}

// the matrix rows of quantity q's IMMERSED nodes from its node table, for
// the current Re / Pe
int tb_ib_rows(tb_ctx *c, int q) {
    const int n = c->nib[q];
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_ib_rows, dim3((n + 63) / 64), dim3(64), 0, c->s, c->d, (const tb_ib_node *)c->ib[q], n,
                       ap_solid_for(c, q), c->ib_row[q], c->ib_cnt[q], c->ib_col[q], c->ib_val[q]);
    TB_LAUNCH_CHECK(c);
    return 0;
}

int tb_geometry_run(tb_ctx *c, const double *h_xz) {
    const Dims &d = c->d;
    Tiles all = make_tiles(d.NX, d.NY, d.NZ);
    c->has_geometry = (h_xz != nullptr);
    c->reinit_its = 0;
    double *sdf_c = c->f[TB_F_SDF_C];
    if (c->has_geometry) {
        double *dh = nullptr, *p0 = c->kv[0], *Dx = c->kv[1], *Dy = c->kv[2], *Dz = c->kv[3], *pold = c->kv[4];
        TB_CHECK(c, hipMalloc((void **)&dh, sizeof(double) * (size_t)d.NX * d.NZ));
        TB_CHECK(c, hipMemcpyAsync(dh, h_xz, sizeof(double) * (size_t)d.NX * d.NZ, hipMemcpyHostToDevice, c->s));
        double offset = c->cfg.bottom_offset ? c->hgrid[13][0] * 2.001 : 0.0;   // Grid.c:913
        hipLaunchKernelGGL(k_sdf_guess, tiles_grid(all), tiles_block(), 0, c->s, d, c->g, dh, offset, p0, sdf_c, all);
        TB_LAUNCH_CHECK(c);
        TB_CHECK(c, hipMemsetAsync(pold, 0, sizeof(double) * d.ntot, c->s));
        // Levelset_compute_fictitious_time (Levelset.c:590-637)
        double dmin_[3] = {1000.0, 1000.0, 1000.0};
        for (int a = 0; a < d.NX - 1; a++) if (c->hgrid[12][a] < dmin_[0]) dmin_[0] = c->hgrid[12][a];
        for (int a = 0; a < d.NY - 1; a++) if (c->hgrid[13][a] < dmin_[1]) dmin_[1] = c->hgrid[13][a];
        for (int a = 0; a < d.NZ - 1; a++) if (c->hgrid[14][a] < dmin_[2]) dmin_[2] = c->hgrid[14][a];
        double dm = dmin_[0] > dmin_[1] ? dmin_[1] : dmin_[0];
        dm = dm > dmin_[2] ? dmin_[2] : dm;
        double dt = 0.5 * dm;
        Tiles inner = make_tiles(d.NX - 1, d.NY - 1, d.NZ - 1);
        double *part = nullptr;
        TB_CHECK(c, hipMalloc((void **)&part, sizeof(double) * (size_t)(all.per_xcd * NXCD)));
        const int max_iterations = 500;   // Surface.c:199
        int converged = 0, it = 0;
        while (!converged) {
            hipLaunchKernelGGL(k_ls_deriv, tiles_grid(all), tiles_block(), 0, c->s, d, c->g, sdf_c, Dx, Dy, Dz,
                               1.0 / c->dEta, 1.0 / c->dXi, 1.0 / c->dPhi, all);
            hipLaunchKernelGGL(k_ls_godunov, tiles_grid(inner), tiles_block(), 0, c->s, d, c->g, p0, sdf_c, Dx, Dy,
                               Dz, dt, c->cfg.reinit_guard == TB_REINIT_SCALED ? 1 : 0, inner);
            hipLaunchKernelGGL(k_ls_norm, tiles_grid(all), tiles_block(), 0, c->s, d, c->g, sdf_c, pold, part, all);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c->s, part, all.per_xcd * NXCD, c->scal + S_TMP0);
            TB_LAUNCH_CHECK(c);
            TB_CHECK(c, hipMemcpyAsync(c->hscal + S_TMP0, c->scal + S_TMP0, sizeof(double), hipMemcpyDeviceToHost, c->s));
            TB_CHECK(c, hipStreamSynchronize(c->s));
            double W_E = sqrt(c->hscal[S_TMP0]);
            int flag = (W_E < 1.0e-5);
            if ((it > max_iterations + 1) || flag) converged = 1;
            it++;
        }
        c->reinit_its = it;
        hipLaunchKernelGGL(k_ls_last_x, dim3((unsigned)(((long)d.NY * d.NZ + 255) / 256)), dim3(256), 0, c->s, d, sdf_c);
        hipLaunchKernelGGL(k_ls_last_y, dim3((unsigned)(((long)d.NX * d.NZ + 255) / 256)), dim3(256), 0, c->s, d, sdf_c);
        hipLaunchKernelGGL(k_ls_last_z, dim3((unsigned)(((long)d.NX * d.NY + 255) / 256)), dim3(256), 0, c->s, d, sdf_c);
        hipLaunchKernelGGL(k_q_sdf, tiles_grid(all), tiles_block(), 0, c->s, d, c->g, sdf_c, c->f[TB_F_SDF_U],
                           c->f[TB_F_SDF_V], c->f[TB_F_SDF_W], all);
        TB_LAUNCH_CHECK(c);
        TB_CHECK(c, hipStreamSynchronize(c->s));
        hipFree(part);
        hipFree(dh);
    } else {
        std::vector<double> ones((size_t)d.ntot, 1.0);
        for (int q = 0; q < 4; q++)
            TB_CHECK(c, hipMemcpyAsync(c->f[TB_F_SDF_U + q], ones.data(), sizeof(double) * d.ntot,
                                       hipMemcpyHostToDevice, c->s));
        TB_CHECK(c, hipStreamSynchronize(c->s));
    }
    // classification
    uint8_t *pre = nullptr;
    TB_CHECK(c, hipMalloc((void **)&pre, (size_t)d.ntot));
    for (int q = 0; q < 4; q++) {
        TB_CHECK(c, hipMemsetAsync(c->st[q], 0, (size_t)d.ntot, c->s));
        hipLaunchKernelGGL(k_classify, tiles_grid(all), tiles_block(), 0, c->s, d, q, (int)c->has_geometry,
                           c->f[TB_F_SDF_U + q], c->st[q], all);
        TB_LAUNCH_CHECK(c);
        if (q != TB_Q_C && c->has_geometry) {
            TB_CHECK(c, hipMemcpyAsync(pre, c->st[q], (size_t)d.ntot, hipMemcpyDeviceToDevice, c->s));
            hipLaunchKernelGGL(k_buffer, tiles_grid(all), tiles_block(), 0, c->s, d, pre, c->st[q], all);
            TB_LAUNCH_CHECK(c);
        }
    }
    TB_CHECK(c, hipStreamSynchronize(c->s));
    hipFree(pre);
    // IB tables: the lexicographic (k, j, i) list of IMMERSED nodes
    // (Immersed_set_q_immersed_indices, Immersed.c:379-445) is compacted on
    // the host; everything per node runs on the device.
    std::vector<uint8_t> hf((size_t)d.NX * d.NY * d.NZ);
    for (int q = 0; q < 4; q++) {
        tb_free_ib(c, q);
        if (!c->has_geometry) continue;
        TB_CHECK(c, hipMemcpy2D(hf.data(), d.NX, c->st[q], d.px, d.NX, (size_t)d.NY * d.NZ, hipMemcpyDeviceToHost));
        std::vector<tb_ib_node> nodes;
        for (int k = 0; k < d.NZ; k++)
            for (int j = 0; j < d.NY; j++)
                for (int i = 0; i < d.NX; i++)
                    if (hf[((size_t)k * d.NY + j) * d.NX + i] == TB_IMMERSED) {
                        tb_ib_node nd;
                        memset(&nd, 0, sizeof(nd));
                        nd.i = i; nd.j = j; nd.k = k;
                        nodes.push_back(nd);
                    }
        int n = (int)nodes.size();
        c->nib[q] = n;
        if (n == 0) continue;
        TB_CHECK(c, hipMalloc((void **)&c->ib[q], sizeof(tb_ib_node) * n));
        TB_CHECK(c, hipMalloc((void **)&c->ib_row[q], sizeof(long) * n));
        TB_CHECK(c, hipMalloc((void **)&c->ib_cnt[q], sizeof(int) * n));
        TB_CHECK(c, hipMalloc((void **)&c->ib_col[q], sizeof(long) * 9 * (size_t)n));
        TB_CHECK(c, hipMalloc((void **)&c->ib_val[q], sizeof(double) * 9 * (size_t)n));
        TB_CHECK(c, hipMemcpyAsync(c->ib[q], nodes.data(), sizeof(tb_ib_node) * n, hipMemcpyHostToDevice, c->s));
        QCoords qc;
        if (q == TB_Q_U) { qc.x = c->g.xu; qc.y = c->g.yu; qc.z = c->g.zu; }
        else if (q == TB_Q_V) { qc.x = c->g.xv; qc.y = c->g.yv; qc.z = c->g.zv; }
        else if (q == TB_Q_W) { qc.x = c->g.xw; qc.y = c->g.yw; qc.z = c->g.zw; }
        else { qc.x = c->g.xc; qc.y = c->g.yc; qc.z = c->g.zc; }
        hipLaunchKernelGGL(k_ib_nodes, dim3((n + 63) / 64), dim3(64), 0, c->s, d, c->g, q, qc,
                           c->f[TB_F_SDF_U + q], c->st[q], c->ib[q], n);
        TB_LAUNCH_CHECK(c);
        if (tb_ib_rows(c, q)) return -1;
        TB_CHECK(c, hipStreamSynchronize(c->s));
    }
    return 0;
}

extern "C" int tb_setup_geometry(tb_ctx *c, const double *h_xz) {
    if (!c) return -1;
    // the increments of earlier Helmholtz solves belong to another operator
    for (int q = 0; q < 4; q++) c->helm[q].inc_ok = false;
    for (int n = 0; n < TB_NCONC_MAX; n++) c->cs[n].h.inc_ok = false;
    return tb_geometry_run(c, h_xz);
}
