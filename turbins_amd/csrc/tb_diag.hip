// tb_diag.hip -- the output-stage diagnostics of the reference
// (Output_flow_properties, Output.c:2062-2216): viscous dissipation,
// kinetic/potential energy conversion, potential energies, bottom wall
// shear, span-averaged height, front location and deposited height.
//
// Sums over cells are block partial sums added in a fixed order (the same
// value on every run).  Per-column quantities are computed one thread per
// (i, k) column with the reference's operations in its order; sums over
// columns are taken on the host in the reference's loop order, so those
// are its bits.
#include <cmath>
#include <cstring>
#include <vector>

#include "tb_common.h"

namespace {

// Grid_set_interface_y_index (Grid.c:1972-2009): the first IMMERSED cell of
// column (i, k) from the bottom, or row 0 when it is FLUID; -1 if neither
__device__ int iface_j(const Dims &d, const uint8_t *__restrict__ sc, int i, int k) {
    for (int j = 0; j < d.NY; j++) {
        const int s = sc[gidx(d, i, j, k)];
        if (s == TB_IMMERSED) return j;
        if (j == 0 && s == TB_FLUID) return j;
    }
    return -1;
}

// ---------------------------------------------------------------------------
// Extract_viscous_dissipation_rate (Extract.c:453-725)
struct ViscArgs {
    Dims d;
    GridDev g;
    const double *u, *v, *w;
    const uint8_t *sc;
    double k_left, k_right, k_bottom, k_top, k_back, k_front;   // ghost factors
    double Cx, Cy, Cz;
    const double *uin, *uout, *vout, *wout;   // planes of open x faces (null: closed)
};

__device__ double visc_cell(const ViscArgs &A, int i, int j, int k) {
    const Dims &d = A.d;
    const GridDev &g = A.g;
    const double *u = A.u, *v = A.v, *w = A.w;
    const long id = gidx(d, i, j, k), rx = d.px, pl = d.plane;
    const int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
    const double Cx = A.Cx, Cy = A.Cy, Cz = A.Cz;
    const double U_ae = Cx * g.mxu[i] * g.mxc[i];
    double U_aw = Cx * g.mxu[i];
    const double U_an = Cy * g.myc[j] * g.myv[j + 1], U_as = Cy * g.myc[j] * g.myv[j];
    const double U_af = Cz * g.mzc[k] * g.mzw[k + 1], U_ab = Cz * g.mzc[k] * g.mzw[k];
    const double V_ae = Cx * g.mxc[i] * g.mxu[i + 1], V_aw = Cx * g.mxc[i] * g.mxu[i];
    const double V_an = Cy * g.myv[j] * g.myc[j];
    double V_as = Cy * g.myv[j];
    const double V_af = Cz * g.mzc[k] * g.mzw[k + 1], V_ab = Cz * g.mzc[k] * g.mzw[k];
    const double W_ae = Cx * g.mxc[i] * g.mxu[i + 1], W_aw = Cx * g.mxc[i] * g.mxu[i];
    const double W_an = Cy * g.myc[j] * g.myv[j + 1], W_as = Cy * g.myc[j] * g.myv[j];
    const double W_af = Cz * g.mzw[k] * g.mzc[k];
    double W_ab = Cz * g.mzw[k];
    const double u_ = u[id], v_ = v[id], w_ = w[id];
    double U_w, V_w, W_w, U_e, V_e, W_e, U_s, V_s, W_s, U_n, V_n, W_n, U_b, V_b, W_b, U_f, V_f, W_f;
    if (i == 0) {
        U_aw *= g.mxc[i];
        U_w = A.uin ? A.uin[(long)k * d.NY + j] : 0.0;   // Extract.c:579-585
        V_w = A.k_left * v_; W_w = A.k_left * w_;
    } else {
        U_aw *= g.mxc[i - 1];
        U_w = u[id - 1]; V_w = v[id - 1]; W_w = w[id - 1];
    }
    if (j == 0) {
        V_as *= g.myc[j];
        U_s = A.k_bottom * u_; V_s = 0.0; W_s = A.k_bottom * w_;
    } else {
        V_as *= g.myc[j - 1];
        U_s = u[id - rx]; V_s = v[id - rx]; W_s = w[id - rx];
    }
    if (k == 0) {
        W_ab *= g.mzc[k];
        U_b = A.k_back * u_; V_b = A.k_back * v_; W_b = 0.0;
    } else {
        W_ab *= g.mzc[k - 1];
        U_b = u[id - pl]; V_b = v[id - pl]; W_b = w[id - pl];
    }
    if (i == NI - 1) {
        if (A.uout) {   // Extract.c:632-639
            const long p = (long)k * d.NY + j;
            U_e = A.uout[p]; V_e = A.vout[p]; W_e = A.wout[p];
        } else {
            U_e = 0.0; V_e = A.k_right * v_; W_e = A.k_right * w_;
        }
    } else {
        U_e = u[id + 1]; V_e = v[id + 1]; W_e = w[id + 1];
    }
    if (j == NJ - 1) {
        U_n = A.k_top * u_; V_n = 0.0; W_n = A.k_top * w_;
    } else {
        U_n = u[id + rx]; V_n = v[id + rx]; W_n = w[id + rx];
    }
    if (k == NK - 1) {
        U_f = A.k_front * u_; V_f = A.k_front * v_; W_f = 0.0;
    } else {
        U_f = u[id + pl]; V_f = v[id + pl]; W_f = w[id + pl];
    }
    const double U_ap = -(U_ae + U_aw + U_ab + U_af + U_as + U_an);
    const double V_ap = -(V_ae + V_aw + V_ab + V_af + V_as + V_an);
    const double W_ap = -(W_ae + W_aw + W_ab + W_af + W_as + W_an);
    const double lu = (U_ae * U_e + U_aw * U_w) + (U_an * U_n + U_as * U_s) + (U_af * U_f + U_ab * U_b) + U_ap * u_;
    const double lv = (V_ae * V_e + V_aw * V_w) + (V_an * V_n + V_as * V_s) + (V_af * V_f + V_ab * V_b) + V_ap * v_;
    const double lw = (W_ae * W_e + W_aw * W_w) + (W_an * W_n + W_as * W_s) + (W_af * W_f + W_ab * W_b) + W_ap * w_;
    const double dV_u = (g.xc[i + 1] - g.xc[i]) * g.dyc[j] * g.dzc[k];
    const double dV_v = (g.yc[j + 1] - g.yc[j]) * g.dxc[i] * g.dzc[k];
    const double dV_w = (g.zc[k + 1] - g.zc[k]) * g.dxc[i] * g.dyc[j];
    return u_ * lu * dV_u + v_ * lv * dV_v + w_ * lw * dV_w;
}

__global__ void __launch_bounds__(256) k_visc(ViscArgs A, double *part, Tiles t) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    int i, j, k;
    if (tile_point(t, i, j, k) && A.sc[gidx(A.d, i, j, k)] == TB_FLUID) acc[0] = visc_cell(A, i, j, k);
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x] = acc[0];
}

// ---------------------------------------------------------------------------
// Conc_compute_potential_energies (Conc.c:1934-2058) and
// Extract_kinetic_potential_conversion_rate (Extract.c:727-802): over FLUID
// cells the sums of c dV (y_c - y_ref), c dV, v_c c dV and -v_c c_total dV
__global__ void __launch_bounds__(256) k_energy(Dims d, GridDev g, const double *__restrict__ v,
                                                const double *__restrict__ c, const double *__restrict__ ct,
                                                const uint8_t *__restrict__ sc, double y_ref, double *part, Tiles t) {
    __shared__ double sh[64];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int i, j, k;
    if (tile_point(t, i, j, k)) {
        const long id = gidx(d, i, j, k);
        if (sc[id] == TB_FLUID) {
            const double dx = g.dxc[i], dy = g.dyc[j], dz = g.dzc[k];
            const double dV = dx * dy * dz;
            const double v_ = interp2d(v[id], g.yv[j], v[id + d.px], g.yv[j + 1], g.yc[j]);
            acc[0] = c[id] * dV * (g.yc[j] - y_ref);
            acc[1] = c[id] * dV;
            acc[2] = v_ * c[id] * (dx * dy * dz);
            acc[3] = -v_ * ct[id] * (dx * dy * dz);
        }
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        for (int a = 0; a < 4; a++) part[a * RED_MAX_BLOCKS + blockIdx.x] = acc[a];
}

// ---------------------------------------------------------------------------
// per-column kernels, one thread per (i, k), a = k * NX + i

// the passive potential energy of a column's deposit (Conc.c:1950-1975)
__global__ void k_col_passive(Dims d, GridDev g, const uint8_t *__restrict__ sc, const double *__restrict__ dep,
                              double y_ref, double *__restrict__ out) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= (long)d.NX * d.NZ) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    const int j = iface_j(d, sc, i, k);
    out[a] = j >= 0 ? dep[a] * g.dxc[i] * g.dzc[k] * (g.yc[j] - y_ref) : 0.0;
}

// Conc_integrate_deposited_height (Conc.c:1675-1749)
__global__ void k_dep_integrate(Dims d, const uint8_t *__restrict__ sc, const double *__restrict__ c,
                                const double *__restrict__ c_old, double vs_abs, double weight, double dt,
                                double *__restrict__ dep) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= (long)d.NX * d.NZ) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    const int j = iface_j(d, sc, i, k);
    if (j < 0) return;
    const long id = gidx(d, i, j, k);
    const double delta_h = dt * vs_abs * weight * 0.5 * (c[id] + c_old[id]);
    dep[a] += delta_h;
}

// Conc_dump_particles (Conc.c:1840-1933): deposit + weight * sum_FLUID c dy
__global__ void k_dep_dumped(Dims d, GridDev g, const uint8_t *__restrict__ sc, const double *__restrict__ c,
                             const double *__restrict__ dep, double weight, double *__restrict__ out) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= (long)d.NX * d.NZ) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    double t = 0.0;
    for (int j = 0; j < d.NY; j++) {
        const long id = gidx(d, i, j, k);
        if (sc[id] == TB_FLUID) t += c[id] * g.dyc[j];
    }
    double h = dep ? dep[a] : 0.0;
    h += t * weight;
    out[a] = h;
}

// Conc_compute_ave_height_x (Conc.c:1524-1608) in two passes: the column
// sums over j of c dy dz, then for each i the sum over k, scaled
__global__ void k_ave_cols(Dims d, GridDev g, const uint8_t *__restrict__ sc, const double *__restrict__ c,
                           double *__restrict__ part) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= (long)d.NX * d.NZ) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    const double dz = g.dzc[k];
    double t = 0.0;
    for (int j = 0; j < d.NY; j++) {
        const long id = gidx(d, i, j, k);
        if (sc[id] == TB_FLUID) t += c[id] * g.dyc[j] * dz;
    }
    part[a] = t;
}

__global__ void k_ave_final(Dims d, const double *__restrict__ part, double coef, double *__restrict__ h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.NX) return;
    double t = 0.0;
    for (int k = 0; k < d.NZ; k++) t += part[(long)k * d.NX + i];
    h[i] = t * coef;
}

// ---------------------------------------------------------------------------
// wall shear

// Grid_get_x/y/z_index (Grid.c:1877-1967) with MyMath_binary_search
// (MyMath.c:384-398): the node left of x among xq[0 .. N-1]; -1 left of the
// range, GVG_INF right of it (the z search uses >=)
constexpr int GVG_INF = 10000000;   // definitions.h:105

__device__ int index_of(const double *xq, double x, int N, bool is_z) {
    if (x < xq[0]) return -1;
    if (is_z ? (x >= xq[N - 1]) : (x > xq[N - 1])) return GVG_INF;
    int lo = 0, hi = N - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (x >= xq[mid]) {
            if (x < xq[mid + 1]) return mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return -1;
}

// MyMath_do_trilinear_inter (MyMath.c:234-360) of a field on coordinates
// (xq, yq, zq); every box is local on one device
__device__ double trilinear(const Dims &d, const double *xq, const double *yq, const double *zq,
                            const double *__restrict__ f, double x, double y, double z) {
    const int NI = d.NX - 1, NJ = d.NY - 1, NK = d.NZ - 1;
    int i0 = index_of(xq, x, NI, false), j0 = index_of(yq, y, NJ, false), k0 = index_of(zq, z, NK, true);
    if (i0 == -1 || j0 == -1 || k0 == -1) {
        i0 = i0 > 0 ? i0 : 0;
        j0 = j0 > 0 ? j0 : 0;
        k0 = k0 > 0 ? k0 : 0;
    }
    if (i0 == GVG_INF || j0 == GVG_INF || k0 == GVG_INF) {
        i0 = i0 < NI - 2 ? i0 : NI - 2;
        j0 = j0 < NJ - 2 ? j0 : NJ - 2;
        k0 = k0 < NK - 2 ? k0 : NK - 2;
    }
    const double x0 = xq[i0], y0 = yq[j0], z0 = zq[k0];
    const double x7 = xq[i0 + 1], y7 = yq[j0 + 1], z7 = zq[k0 + 1];
    const double a = (x7 - x) / (x7 - x0), b = (y7 - y) / (y7 - y0), gm = (z7 - z) / (z7 - z0);
    double wc[8];
    wc[0] = a * b * gm;
    wc[1] = (1.0 - a) * b * gm;
    wc[2] = a * (1.0 - b) * gm;
    wc[3] = (1.0 - a) * (1.0 - b) * gm;
    wc[4] = a * b * (1.0 - gm);
    wc[5] = (1.0 - a) * b * (1.0 - gm);
    wc[6] = a * (1.0 - b) * (1.0 - gm);
    wc[7] = (1.0 - a) * (1.0 - b) * (1.0 - gm);
    double val = 0.0;
    for (int m = 0; m < 8; m++) val += wc[m] * f[gidx(d, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + (m >> 2))];
    return val;
}

__device__ double dist3(const double *p, const double *q) {
    // MyMath_get_point_point_distance (MyMath.c:105-111)
    const double d2 = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
    return sqrt(d2);
}

// MyMath_GaussJordan_matrix_inv (MyMath.c:117-209), full pivoting, for the
// 3 x 3 system of the quadratic wall fit
__device__ void gauss_jordan3(double A[3][3]) {
    int index_col[3], index_row[3], index_piv[3] = {-1, -1, -1};
    int in_row = 0, in_col = 0;
    for (int i = 0; i < 3; i++) {
        double mx = 0.0;
        for (int j = 0; j < 3; j++)
            if (index_piv[j] != 0)
                for (int k = 0; k < 3; k++)
                    if (index_piv[k] == -1 && fabs(A[j][k]) >= mx) {
                        mx = fabs(A[j][k]);
                        in_row = j;
                        in_col = k;
                    }
        ++(index_piv[in_col]);
        if (in_row != in_col)
            for (int h = 0; h < 3; h++) {
                const double t = A[in_row][h];
                A[in_row][h] = A[in_col][h];
                A[in_col][h] = t;
            }
        index_row[i] = in_row;
        index_col[i] = in_col;
        const double a_piv = 1.0 / A[in_col][in_col];
        A[in_col][in_col] = 1.0;
        for (int h = 0; h < 3; h++) A[in_col][h] *= a_piv;
        for (int hh = 0; hh < 3; hh++)
            if (hh != in_col) {
                const double t = A[hh][in_col];
                A[hh][in_col] = 0.0;
                for (int h = 0; h < 3; h++) A[hh][h] -= A[in_col][h] * t;
            }
    }
    for (int h = 2; h >= 0; h--)
        if (index_row[h] != index_col[h])
            for (int k = 0; k < 3; k++) {
                const double t = A[k][index_row[h]];
                A[k][index_row[h]] = A[k][index_col[h]];
                A[k][index_col[h]] = t;
            }
}

struct ShearArgs {
    Dims d;
    GridDev g;
    const double *u, *v, *w;
    const uint8_t *sc;
    const tb_ib_node *ib;       // IMMERSED nodes of the concentration grid, (k, j, i) order
    const long *ib_row;         // their flat indices, ascending
    int nib;
    double re;
    int quadratic;
    long ncol;
    double *out;                // tau, u_tan, v_tan, w_tan, ncol each
};

// the velocity at p and its part tangential to n; the normal component is
// signed (Extract.c:125-133) or its magnitude (Extract.c:298-303)
__device__ void tangential(const ShearArgs &A, const double *p, const double *n, bool abs_normal, double t[3]) {
    const GridDev &g = A.g;
    const double ui = trilinear(A.d, g.xu, g.yu, g.zu, A.u, p[0], p[1], p[2]);
    const double vi = trilinear(A.d, g.xv, g.yv, g.zv, A.v, p[0], p[1], p[2]);
    const double wi = trilinear(A.d, g.xw, g.yw, g.zw, A.w, p[0], p[1], p[2]);
    double Vn = ui * n[0] + vi * n[1] + wi * n[2];
    if (abs_normal) Vn = fabs(Vn);
    t[0] = ui - Vn * n[0];
    t[1] = vi - Vn * n[1];
    t[2] = wi - Vn * n[2];
}

__global__ void k_wall_shear(ShearArgs A) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const Dims &d = A.d;
    const GridDev &g = A.g;
    if (a >= A.ncol) return;
    const int i = (int)(a % d.NX), k = (int)(a / d.NX);
    const int j = iface_j(d, A.sc, i, k);
    double t[3] = {0.0, 0.0, 0.0}, tau = 0.0;
    if (j >= 0) {
        const long id = gidx(d, i, j, k);
        const int s = A.sc[id];
        double dist = 0.0, dVt_dn = 0.0;
        if (s == TB_IMMERSED) {
            int lo = 0, hi = A.nib - 1, m = -1;   // the node's entry in the table
            while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                const long r = A.ib_row[mid];
                if (r == id) { m = mid; break; }
                if (r < id) lo = mid + 1;
                else hi = mid - 1;
            }
            if (m >= 0) {
                const tb_ib_node &nd = A.ib[m];
                const double Dt = sqrt(3.0) * g.dyc[nd.j];
                double p1[3], p2[3];
                for (int e = 0; e < 3; e++) p1[e] = nd.control[e] + nd.n[e] * Dt;
                dist = dist3(p1, nd.control);
                tangential(A, p1, nd.n, A.quadratic != 0, t);
                if (A.quadratic) {
                    const double Vt1 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    for (int e = 0; e < 3; e++) p2[e] = nd.control[e] + nd.n[e] * 2.0 * Dt;
                    const double d2 = dist3(p2, nd.control);
                    tangential(A, p2, nd.n, true, t);
                    const double Vt2 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    const double d0 = 0.0, Vt0 = 0.0;
                    double M[3][3] = {{dist * dist, dist, 1.0}, {d2 * d2, d2, 1.0}, {d0 * d0, d0, 1.0}};
                    gauss_jordan3(M);
                    const double qa = Vt1 * M[0][0] + Vt2 * M[0][1] + Vt0 * M[0][2];
                    const double qb = Vt1 * M[1][0] + Vt2 * M[1][1] + Vt0 * M[1][2];
                    dVt_dn = 2.0 * qa * d0 + qb;
                }
            }
        }
        if (j == 0 && s == TB_FLUID) {
            t[0] = interp2d(A.u[id], g.xu[i], A.u[id + 1], g.xu[i + 1], g.xc[i]);
            t[1] = 0.0;
            t[2] = interp2d(A.w[id], g.zw[k], A.w[id + d.plane], g.zw[k + 1], g.zc[k]);
            dist = g.yc[j] - g.yv[j];
            if (A.quadratic) dVt_dn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * (1.0 / dist);
        }
        if (A.quadratic) tau = dVt_dn * (1.0 / A.re);
        else tau = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) / (dist * A.re);
    }
    A.out[a] = tau;
    A.out[A.ncol + a] = t[0];
    A.out[2 * A.ncol + a] = t[1];
    A.out[3 * A.ncol + a] = t[2];
}

// ---------------------------------------------------------------------------
// host side

int cell_sums(tb_ctx *c, bool visc, int nslots, double *out, const char *who) {
    const Tiles t = make_tiles(c->ncx, c->ncy, c->ncz);
    const dim3 grid = tiles_grid(t);
    if ((long)grid.x > RED_MAX_BLOCKS) {
        c->err = std::string(who) + ": grid too large for the reduction buffer";
        return -1;
    }
    if (visc) {
        ViscArgs A;
        A.d = c->d;
        A.g = c->g;
        A.u = c->f[TB_F_U]; A.v = c->f[TB_F_V]; A.w = c->f[TB_F_W];
        A.sc = c->st[TB_Q_C];
        const int *wl = c->cfg.wall_noslip;   // Boundary.h: no-slip -1, free-slip +1 (Extract.c:487-522)
        A.k_left = wl[0] ? -1.0 : 1.0;
        A.k_right = wl[1] ? -1.0 : 1.0;
        A.k_bottom = wl[2] ? -1.0 : 1.0;
        A.k_top = wl[3] ? -1.0 : 1.0;
        A.k_back = wl[4] ? -1.0 : 1.0;
        A.k_front = wl[5] ? -1.0 : 1.0;
        A.Cx = 1.0 / (c->dEta * c->dEta);
        A.Cy = 1.0 / (c->dXi * c->dXi);
        A.Cz = 1.0 / (c->dPhi * c->dPhi);
        A.uin = tb_in_plane(c, TB_Q_U);
        A.uout = tb_out_plane(c, TB_Q_U);
        A.vout = tb_out_plane(c, TB_Q_V);
        A.wout = tb_out_plane(c, TB_Q_W);
        hipLaunchKernelGGL(k_visc, grid, tiles_block(), 0, c->s, A, c->red, t);
    } else {
        if (tb_total_concentration(c)) return -1;
        const double *ct = tb_buoyancy_source(c);
        hipLaunchKernelGGL(k_energy, grid, tiles_block(), 0, c->s, c->d, c->g, (const double *)c->f[TB_F_V],
                           (const double *)c->f[TB_F_C], ct, (const uint8_t *)c->st[TB_Q_C], c->hgrid[7][0],
                           c->red, t);
    }
    TB_LAUNCH_CHECK(c);
    if (tb_reduce_finish(c, (int)grid.x, nslots, S_TMP0)) return -1;
    if (tb_read_scalars(c, S_TMP0, nslots)) return -1;
    for (int a = 0; a < nslots; a++) out[a] = c->hscal[S_TMP0 + a];
    return 0;
}

int column_values(tb_ctx *c, std::vector<double> &h, double *dev) {
    const long n = (long)c->d.NX * c->d.NZ;
    h.resize((size_t)n);
    TB_CHECK(c, hipMemcpyAsync(h.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

double *deposit_of(tb_ctx *c, bool create) {
    ConcSlot &s = c->cs[c->conc_sel];
    if (!s.dep && create) {
        const size_t bytes = sizeof(double) * (size_t)c->d.NX * c->d.NZ;
        if (hipMalloc((void **)&s.dep, bytes) != hipSuccess) {
            s.dep = nullptr;
            c->err = "deposit: out of device memory";
            return nullptr;
        }
        c->bytes += (int64_t)bytes;
        if (hipMemsetAsync(s.dep, 0, bytes, c->s) != hipSuccess) return nullptr;
    }
    return s.dep;
}

dim3 column_grid(const tb_ctx *c) { return dim3((unsigned)(((long)c->d.NX * c->d.NZ + 255) / 256)); }

// the selected field's v_settl0 (0 for a non-particle field, Conc.c:169-185)
double field_vs_abs(const tb_ctx *c) { return c->cfg.conc_type == TB_PARTICLE ? fabs(c->cfg.v_settle) : 0.0; }

}  // namespace

void tb_free_deposits(tb_ctx *c, int first) {
    for (int n = first; n < TB_NCONC_MAX; n++)
        if (c->cs[n].dep) {
            hipFree(c->cs[n].dep);
            c->cs[n].dep = nullptr;
            c->bytes -= (int64_t)(sizeof(double) * (size_t)c->d.NX * c->d.NZ);
        }
}

extern "C" int tb_viscous_dissipation(tb_ctx *c, double *rate) {
    if (!c || !rate) {
        if (c) c->err = "tb_viscous_dissipation: bad argument";
        return -1;
    }
    double s = 0.0;
    if (cell_sums(c, true, 1, &s, "tb_viscous_dissipation")) return -1;
    s *= 1.0 / c->cfg.re;
    *rate = s;
    return 0;
}

extern "C" int tb_kinetic_potential_conversion(tb_ctx *c, double *rate) {
    if (!c || !rate) {
        if (c) c->err = "tb_kinetic_potential_conversion: bad argument";
        return -1;
    }
    double s[4];
    if (cell_sums(c, false, 4, s, "tb_kinetic_potential_conversion")) return -1;
    *rate = s[3];
    return 0;
}

extern "C" int tb_potential_energies(tb_ctx *c, tb_energies *out) {
    if (!c || !out) {
        if (c) c->err = "tb_potential_energies: bad argument";
        return -1;
    }
    const double y_ref = c->hgrid[7][0];
    const double vs = field_vs_abs(c);
    const double alpha = c->cs[c->conc_sel].p.alpha;
    double ep_passive = 0.0;
    const double *dep = c->cs[c->conc_sel].dep;
    if (c->cfg.conc_type == TB_PARTICLE && dep) {
        double *dev = c->kv[7];
        hipLaunchKernelGGL(k_col_passive, column_grid(c), dim3(256), 0, c->s, c->d, c->g,
                           (const uint8_t *)c->st[TB_Q_C], dep, y_ref, dev);
        TB_LAUNCH_CHECK(c);
        std::vector<double> h;
        if (column_values(c, h, dev)) return -1;
        for (double x : h) ep_passive += x;   // k outer, i inner (Conc.c:1958-1975)
    }
    double s[4];
    if (cell_sums(c, false, 4, s, "tb_potential_energies")) return -1;
    double ep = s[0], stokes = s[1], conv = s[2];
    ep *= alpha;
    stokes *= vs * alpha;
    conv *= alpha;
    out->ep = ep;
    out->ep_passive = ep_passive;
    out->stokes = stokes;
    out->conversion = conv;
    out->ep_passive_rate = ep_passive * vs;
    out->ep_rate = conv + out->ep_passive_rate - stokes;
    return 0;
}

extern "C" int tb_wall_shear(tb_ctx *c, int quadratic, double *tau, double *u_tan, double *v_tan, double *w_tan) {
    if (!c) return -1;
    ShearArgs A;
    A.d = c->d;
    A.g = c->g;
    A.u = c->f[TB_F_U]; A.v = c->f[TB_F_V]; A.w = c->f[TB_F_W];
    A.sc = c->st[TB_Q_C];
    A.ib = c->ib[TB_Q_C];
    A.ib_row = c->ib_row[TB_Q_C];
    A.nib = c->nib[TB_Q_C];
    A.re = c->cfg.re;
    A.quadratic = quadratic != 0;
    A.ncol = (long)c->d.NX * c->d.NZ;
    A.out = c->kv[7];   // 4 column arrays fit: ncol * 4 < (nx+1)(ny+1)(nz+1) for ny >= 4
    hipLaunchKernelGGL(k_wall_shear, column_grid(c), dim3(256), 0, c->s, A);
    TB_LAUNCH_CHECK(c);
    double *dst[4] = {tau, u_tan, v_tan, w_tan};
    for (int a = 0; a < 4; a++)
        if (dst[a])
            TB_CHECK(c, hipMemcpyAsync(dst[a], A.out + a * A.ncol, sizeof(double) * A.ncol, hipMemcpyDeviceToHost,
                                       c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_ave_height_x(tb_ctx *c, double *h) {
    if (!c || !h) {
        if (c) c->err = "tb_ave_height_x: bad argument";
        return -1;
    }
    double *part = c->kv[7], *dh = c->kv[7] + (long)c->d.NX * c->d.NZ;
    hipLaunchKernelGGL(k_ave_cols, column_grid(c), dim3(256), 0, c->s, c->d, c->g, (const uint8_t *)c->st[TB_Q_C],
                       (const double *)c->f[TB_F_C], part);
    hipLaunchKernelGGL(k_ave_final, dim3((unsigned)((c->d.NX + 255) / 256)), dim3(256), 0, c->s, c->d,
                       (const double *)part, 1.0 / (c->cfg.ly * c->cfg.lz), dh);
    TB_LAUNCH_CHECK(c);
    TB_CHECK(c, hipMemcpyAsync(h, dh, sizeof(double) * c->d.NX, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_front_location(tb_ctx *c, double front_limit, double *x_front) {
    if (!c || !x_front) {
        if (c) c->err = "tb_front_location: bad argument";
        return -1;
    }
    const int NX = c->d.NX, sel = c->conc_sel;
    std::vector<double> h((size_t)NX * c->nconc);
    for (int n = 0; n < c->nconc; n++)
        if (tb_conc_use(c, n) || tb_ave_height_x(c, h.data() + (size_t)n * NX)) {
            tb_conc_use(c, sel);
            return -1;
        }
    if (tb_conc_use(c, sel)) return -1;
    const std::vector<double> &xc = c->hgrid[0];
    double front = -1.0;
    for (int i = NX - 2; i > 0; i--) {
        double t0 = 0.0, t1 = 0.0;
        for (int n = 0; n < c->nconc; n++) {
            t0 += h[(size_t)n * NX + i - 1] * c->cs[n].p.alpha;
            t1 += h[(size_t)n * NX + i] * c->cs[n].p.alpha;
        }
        if (t1 <= front_limit && t0 >= front_limit) {
            front = interp2d(xc[i - 1], t0, xc[i], t1, front_limit);
            break;
        }
    }
    *x_front = front;
    return 0;
}

extern "C" int tb_deposit_integrate(tb_ctx *c, double weight, double dt) {
    if (!c) return -1;
    double *dep = deposit_of(c, true);
    if (!dep) return -1;
    hipLaunchKernelGGL(k_dep_integrate, column_grid(c), dim3(256), 0, c->s, c->d, (const uint8_t *)c->st[TB_Q_C],
                       (const double *)c->f[TB_F_C], (const double *)c->old[TB_Q_C], field_vs_abs(c), weight, dt,
                       dep);
    TB_LAUNCH_CHECK(c);
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_deposit_height(tb_ctx *c, double *h) {
    if (!c || !h) {
        if (c) c->err = "tb_deposit_height: bad argument";
        return -1;
    }
    const size_t n = (size_t)c->d.NX * c->d.NZ;
    const double *dep = c->cs[c->conc_sel].dep;
    if (!dep) {
        memset(h, 0, sizeof(double) * n);
        return 0;
    }
    TB_CHECK(c, hipMemcpyAsync(h, dep, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_deposit_set(tb_ctx *c, const double *h) {
    if (!c || !h) {
        if (c) c->err = "tb_deposit_set: bad argument";
        return -1;
    }
    double *dep = deposit_of(c, true);
    if (!dep) return -1;
    TB_CHECK(c, hipMemcpyAsync(dep, h, sizeof(double) * (size_t)c->d.NX * c->d.NZ, hipMemcpyHostToDevice, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_deposit_dumped(tb_ctx *c, double weight, double *h) {
    if (!c || !h) {
        if (c) c->err = "tb_deposit_dumped: bad argument";
        return -1;
    }
    double *dev = c->kv[7];
    hipLaunchKernelGGL(k_dep_dumped, column_grid(c), dim3(256), 0, c->s, c->d, c->g, (const uint8_t *)c->st[TB_Q_C],
                       (const double *)c->f[TB_F_C], (const double *)c->cs[c->conc_sel].dep, weight, dev);
    TB_LAUNCH_CHECK(c);
    TB_CHECK(c, hipMemcpyAsync(h, dev, sizeof(double) * (size_t)c->d.NX * c->d.NZ, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}
