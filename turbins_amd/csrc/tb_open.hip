// tb_open.hip -- the open x faces: inflow profiles (Inflow.c:13-260), the
// convective outflow condition with the global mass correction
// (Outflow.c:12-304), the right-hand-side terms of the rows next to an
// open face (Velocity.c:1691-1721, 1885-1897, 2011-2023; Conc.c:938-963) and
// the outflow planes' RK bookkeeping (Velocity.c:2058-2120,
// Conc.c:1223-1289).  The operators that read the planes inside their own
// kernels (ENO ghosts, transposes, cell centres, Helmholtz rows, divergence,
// dissipation) take them from tb_in_plane / tb_out_plane.
//
// A plane is NY*NZ doubles indexed [k][j], j fastest: entry (j, k) belongs
// to the row of nodes (., j, k) of the padded layout.
#include <cmath>

#include "tb_common.h"

constexpr int OPEN_BLOCK = 1024;   // the one-workgroup reductions

static long plane_n(const tb_ctx *c) { return (long)c->d.NY * c->d.NZ; }

const double *tb_in_plane(const tb_ctx *c, int q) {
    if (!c->obc.inflow || !c->u_in) return nullptr;
    if (q == TB_Q_U) return c->u_in;
    if (q == TB_Q_C) return c->cs[c->conc_sel].cin;
    return nullptr;
}

const double *tb_out_plane(const tb_ctx *c, int q) {
    if (!c->obc.outflow || !c->outp[0]) return nullptr;
    return q < 3 ? c->outp[q] : c->cs[c->conc_sel].cout;
}

void tb_open_free(tb_ctx *c) {
    double **all[] = {&c->u_in, &c->outp[0], &c->outp[1], &c->outp[2], &c->outp_old[0], &c->outp_old[1],
                      &c->outp_old[2], &c->open_scal};
    for (double **p : all) {
        hipFree(*p);
        *p = nullptr;
    }
    for (int n = 0; n < TB_NCONC_MAX; n++) {
        ConcSlot &s = c->cs[n];
        hipFree(s.cin); hipFree(s.cout); hipFree(s.cout_old);
        s.cin = s.cout = s.cout_old = nullptr;
    }
}

__device__ inline bool open_vel_row(int s) { return s == TB_FLUID || s == TB_BUFFER_SOLID; }

// u_in = 1 * tf on the FLUID u nodes of i = 1 (Inflow.c:85-139)
__global__ void k_inflow_u(Dims d, const uint8_t *__restrict__ su, double *__restrict__ uin, double tf) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= (long)d.NY * d.NZ) return;
    const int j = (int)(p % d.NY), k = (int)(p / d.NY);
    if (su[gidx(d, 1, j, k)] == TB_FLUID) {
        double v = 1.0;
        v *= tf;
        uin[p] = v;
    }
}

// c_in = 1 * tf on the non-SOLID cells of i = 0 (Inflow.c:219-252)
__global__ void k_inflow_c(Dims d, const uint8_t *__restrict__ sc, double *__restrict__ cin, double tf) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= (long)d.NY * d.NZ) return;
    const int j = (int)(p % d.NY), k = (int)(p / d.NY);
    if (sc[gidx(d, 0, j, k)] != TB_SOLID) {
        double v = 1.0;
        v *= tf;
        cin[p] = v;
    }
}

// influx = sum of u_in dy dz over the FLUID u nodes of i = 1 (one workgroup)
__global__ void __launch_bounds__(OPEN_BLOCK) k_influx(Dims d, GridDev g, const uint8_t *__restrict__ su,
                                                       const double *__restrict__ uin, double *__restrict__ out) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    const long n = (long)d.NY * d.NZ;
    for (long p = threadIdx.x; p < n; p += OPEN_BLOCK) {
        const int j = (int)(p % d.NY), k = (int)(p / d.NY);
        if (su[gidx(d, 1, j, k)] == TB_FLUID) {
            const double dy = g.yv[j + 1] - g.yv[j];
            const double dz = g.zw[k + 1] - g.zw[k];
            const double dA = dy * dz;
            acc[0] += uin[p] * dA;
        }
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) out[0] = acc[0];
}

// out = (1 - coef) out + coef q[NI-1] where q's node at i = NI-1 is FLUID or
// IMMERSED (Outflow.c:75-83, 146-152)
__global__ void k_outflow(Dims d, const uint8_t *__restrict__ st, const double *__restrict__ q,
                          double *__restrict__ out, double coef) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= (long)d.NY * d.NZ) return;
    const int j = (int)(p % d.NY), k = (int)(p / d.NY);
    const long id = gidx(d, d.NX - 2, j, k);
    const int s = st[id];
    if (s == TB_FLUID || s == TB_IMMERSED) out[p] = (1.0 - coef) * out[p] + coef * q[id];
}

// u's plane: the convective update, then Outflow_update_u_velocity_to_conserve_mass
// (Outflow.c:163-290): outflux and area over the open rows, and the uniform
// correction (influx - outflux) / area on them; one workgroup, each thread
// owning the entries p = t mod OPEN_BLOCK throughout
__global__ void __launch_bounds__(OPEN_BLOCK) k_outflow_u(Dims d, GridDev g, const uint8_t *__restrict__ su,
                                                          const double *__restrict__ u, double *__restrict__ out,
                                                          double coef, const double *__restrict__ influx) {
    __shared__ double sh[32];
    __shared__ double corr;
    double acc[2] = {0.0, 0.0};
    const long n = (long)d.NY * d.NZ;
    for (long p = threadIdx.x; p < n; p += OPEN_BLOCK) {
        const int j = (int)(p % d.NY), k = (int)(p / d.NY);
        const long id = gidx(d, d.NX - 2, j, k);
        const int s = su[id];
        if (s == TB_FLUID || s == TB_IMMERSED) {
            const double o = (1.0 - coef) * out[p] + coef * u[id];
            out[p] = o;
            const double dy = g.yv[j + 1] - g.yv[j];
            const double dz = g.zw[k + 1] - g.zw[k];
            const double dA = dy * dz;
            acc[0] += dA;
            acc[1] += dA * o;
        }
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) corr = acc[0] > 0.0 ? (influx[0] - acc[1]) / acc[0] : 0.0;
    __syncthreads();
    const double cr = corr;
    for (long p = threadIdx.x; p < n; p += OPEN_BLOCK) {
        const int j = (int)(p % d.NY), k = (int)(p / d.NY);
        const int s = su[gidx(d, d.NX - 2, j, k)];
        if (s == TB_FLUID || s == TB_IMMERSED) out[p] += cr;
    }
}

// the plane value times the coefficient the open face dropped from the row,
// added to the right-hand side: u at i = 1 (inflow) and i = NX-2 (outflow),
// v and w at i = NX-2, c at i = 0 and NX-2 (whose outflow coefficient uses
// metric_xu[i], Conc.c:957)
__global__ void k_open_rhs(Dims d, GridDev g, int q, const uint8_t *__restrict__ st, double *__restrict__ rhs,
                           const double *__restrict__ pin, const double *__restrict__ pout, double Cx) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= (long)d.NY * d.NZ) return;
    const int j = (int)(p % d.NY), k = (int)(p / d.NY);
    const int il = d.NX - 2;
    if (q == TB_Q_U) {
        if (pin) {
            const long id = gidx(d, 1, j, k);
            if (open_vel_row(st[id])) {
                const double aw = -Cx * g.mxu[1] * g.mxc[0];
                const double value = aw * pin[p];
                rhs[id] = rhs[id] + value;
            }
        }
        if (pout) {
            const long id = gidx(d, il, j, k);
            if (open_vel_row(st[id])) {
                const double ae = -Cx * g.mxu[il] * g.mxc[il];
                const double value = ae * pout[p];
                rhs[id] = rhs[id] + value;
            }
        }
    } else if (q == TB_Q_C) {
        if (pin) {
            const long id = gidx(d, 0, j, k);
            if (st[id] == TB_FLUID) {
                const double aw = -Cx * g.mxc[0] * g.mxu[0];
                const double value = aw * pin[p];
                rhs[id] = rhs[id] + value;
            }
        }
        if (pout) {
            const long id = gidx(d, il, j, k);
            if (st[id] == TB_FLUID) {
                const double ae = -Cx * g.mxc[il] * g.mxu[il];
                const double value = ae * pout[p];
                rhs[id] = rhs[id] + value;
            }
        }
    } else if (pout) {
        const long id = gidx(d, il, j, k);
        if (open_vel_row(st[id])) {
            const double ae = -Cx * g.mxc[il] * g.mxu[il + 1];
            const double value = ae * pout[p];
            rhs[id] = rhs[id] + value;
        }
    }
}

__global__ void k_plane_copy(double *__restrict__ dst, const double *__restrict__ src, long n) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

// VecAXPBY order of Velocity_rk_average_outflow: old_frac old + new_frac out
__global__ void k_plane_rk(double *__restrict__ out, const double *__restrict__ old, double a, double b, long n) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < n) out[p] = a * old[p] + b * out[p];
}

static dim3 plane_grid(const tb_ctx *c) { return dim3((unsigned)((plane_n(c) + 255) / 256)); }

static int plane_alloc(tb_ctx *c, double **p) {
    if (*p) return 0;
    const size_t bytes = sizeof(double) * (size_t)plane_n(c);
    TB_CHECK(c, hipMalloc((void **)p, bytes));
    c->bytes += (int64_t)bytes;
    return 0;
}

extern "C" int tb_set_open_boundaries(tb_ctx *c, const tb_open_bc *bc) {
    if (!c || !bc || (bc->inflow != 0 && bc->inflow != 1) || (bc->outflow != 0 && bc->outflow != 1) ||
        !std::isfinite(bc->u_conv) || (bc->transient_inflow && !(bc->time_c > 0.0))) {
        if (c) c->err = "tb_set_open_boundaries: bad argument";
        return -1;
    }
    if (plane_alloc(c, &c->u_in) || plane_alloc(c, &c->open_scal)) return -1;
    for (int q = 0; q < 3; q++)
        if (plane_alloc(c, &c->outp[q]) || plane_alloc(c, &c->outp_old[q])) return -1;
    for (int n = 0; n < TB_NCONC_MAX; n++) {
        ConcSlot &s = c->cs[n];
        if (plane_alloc(c, &s.cin) || plane_alloc(c, &s.cout) || plane_alloc(c, &s.cout_old)) return -1;
    }
    // planes and influx start at zero, as Memory_allocate_2D_double's arrays
    const size_t bytes = sizeof(double) * (size_t)plane_n(c);
    double *all[] = {c->u_in, c->open_scal, c->outp[0], c->outp[1], c->outp[2], c->outp_old[0], c->outp_old[1],
                     c->outp_old[2]};
    for (double *p : all) TB_CHECK(c, hipMemsetAsync(p, 0, bytes, c->s));
    for (int n = 0; n < TB_NCONC_MAX; n++) {
        ConcSlot &s = c->cs[n];
        TB_CHECK(c, hipMemsetAsync(s.cin, 0, bytes, c->s));
        TB_CHECK(c, hipMemsetAsync(s.cout, 0, bytes, c->s));
        TB_CHECK(c, hipMemsetAsync(s.cout_old, 0, bytes, c->s));
    }
    c->obc = *bc;
    // the rows next to the faces change: earlier increments are no guess
    tb_conc_sync(c);
    for (int q = 0; q < 4; q++) c->helm[q].inc_ok = false;
    for (int n = 0; n < TB_NCONC_MAX; n++) c->cs[n].h.inc_ok = false;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

static int need_planes(tb_ctx *c, const char *who) {
    if (c->u_in) return 0;
    c->err = std::string(who) + ": tb_set_open_boundaries was not called";
    return -1;
}

static int read_influx(tb_ctx *c, double *influx) {
    if (!influx) return 0;
    TB_CHECK(c, hipMemcpyAsync(influx, c->open_scal, sizeof(double), hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_inflow_profile(tb_ctx *c, double time, double *influx) {
    if (!c) return -1;
    if (need_planes(c, "tb_inflow_profile")) return -1;
    const double tf = c->obc.transient_inflow ? (1.0 - exp(-(time + 0.01) / (c->obc.time_c))) : 1.0;
    hipLaunchKernelGGL(k_inflow_u, plane_grid(c), dim3(256), 0, c->s, c->d, (const uint8_t *)c->st[0], c->u_in, tf);
    hipLaunchKernelGGL(k_influx, dim3(1), dim3(OPEN_BLOCK), 0, c->s, c->d, c->g, (const uint8_t *)c->st[0],
                       (const double *)c->u_in, c->open_scal);
    for (int n = 0; n < c->nconc; n++)
        hipLaunchKernelGGL(k_inflow_c, plane_grid(c), dim3(256), 0, c->s, c->d, (const uint8_t *)c->st[3],
                           c->cs[n].cin, tf);
    TB_LAUNCH_CHECK(c);
    if (read_influx(c, influx)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

static double *plane_of(tb_ctx *c, int which, int q) {
    if (q < 0 || q > 3) return nullptr;
    if (which == TB_PLANE_INFLOW) {
        if (q == TB_Q_U) return c->u_in;
        if (q == TB_Q_C) return c->cs[c->conc_sel].cin;
        return nullptr;
    }
    if (which == TB_PLANE_OUTFLOW) return q < 3 ? c->outp[q] : c->cs[c->conc_sel].cout;
    return nullptr;
}

extern "C" int tb_set_boundary_plane(tb_ctx *c, int which, int q, const double *plane) {
    if (!c) return -1;
    if (need_planes(c, "tb_set_boundary_plane")) return -1;
    double *p = plane_of(c, which, q);
    if (!p || !plane) {
        c->err = "tb_set_boundary_plane: no such plane";
        return -1;
    }
    TB_CHECK(c, hipMemcpyAsync(p, plane, sizeof(double) * (size_t)plane_n(c), hipMemcpyHostToDevice, c->s));
    if (which == TB_PLANE_INFLOW && q == TB_Q_U)
        hipLaunchKernelGGL(k_influx, dim3(1), dim3(OPEN_BLOCK), 0, c->s, c->d, c->g, (const uint8_t *)c->st[0],
                           (const double *)c->u_in, c->open_scal);
    TB_LAUNCH_CHECK(c);
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_get_boundary_plane(tb_ctx *c, int which, int q, double *plane) {
    if (!c) return -1;
    if (need_planes(c, "tb_get_boundary_plane")) return -1;
    double *p = plane_of(c, which, q);
    if (!p || !plane) {
        c->err = "tb_get_boundary_plane: no such plane";
        return -1;
    }
    TB_CHECK(c, hipMemcpyAsync(plane, p, sizeof(double) * (size_t)plane_n(c), hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_get_influx(tb_ctx *c, double *influx) {
    if (!c || !influx) return -1;
    if (need_planes(c, "tb_get_influx")) return -1;
    return read_influx(c, influx);
}

// Outflow_impose_convective_boundary, queued on the context's stream
int tb_outflow_async(tb_ctx *c, double dt) {
    const int NI = c->d.NX - 1;
    const double dx = c->hgrid[3][NI] - c->hgrid[3][NI - 1];   // xu[NI] - xu[NI-1]
    const double coef = dt * c->obc.u_conv / dx;
    tb_conc_sync(c);
    hipLaunchKernelGGL(k_outflow_u, dim3(1), dim3(OPEN_BLOCK), 0, c->s, c->d, c->g, (const uint8_t *)c->st[0],
                       (const double *)c->f[TB_F_U], c->outp[0], coef, (const double *)c->open_scal);
    for (int q = 1; q < 3; q++)
        hipLaunchKernelGGL(k_outflow, plane_grid(c), dim3(256), 0, c->s, c->d, (const uint8_t *)c->st[q],
                           (const double *)c->f[TB_F_U + q], c->outp[q], coef);
    for (int n = 0; n < c->nconc; n++)
        hipLaunchKernelGGL(k_outflow, plane_grid(c), dim3(256), 0, c->s, c->d, (const uint8_t *)c->st[3],
                           (const double *)c->cs[n].c, c->cs[n].cout, coef);
    TB_LAUNCH_CHECK(c);
    return 0;
}

extern "C" int tb_outflow_convective(tb_ctx *c, double dt) {
    if (!c || !(dt > 0.0)) { if (c) c->err = "tb_outflow_convective: bad argument"; return -1; }
    if (need_planes(c, "tb_outflow_convective")) return -1;
    if (!c->obc.outflow) {
        c->err = "tb_outflow_convective: the outflow face is closed";
        return -1;
    }
    if (tb_outflow_async(c, dt)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

// q = -1: u, v, w and every concentration field; q = 3: the selected field
int tb_open_rhs(tb_ctx *c, int qsel) {
    if (!c->obc.inflow && !c->obc.outflow) return 0;
    tb_conc_sync(c);
    const double re = c->cfg.re;
    for (int q = 0; q < 3; q++) {
        if (qsel >= 0 && q != qsel) continue;
        const double *pin = q == TB_Q_U ? tb_in_plane(c, q) : nullptr;
        const double *pout = tb_out_plane(c, q);
        if (!pin && !pout) continue;
        const double Cx = -1.0 / (re * c->dEta * c->dEta);
        hipLaunchKernelGGL(k_open_rhs, plane_grid(c), dim3(256), 0, c->s, c->d, c->g, q, (const uint8_t *)c->st[q],
                           c->f[TB_F_RHS_U + q], pin, pout, Cx);
    }
    for (int n = 0; n < c->nconc; n++) {
        if (qsel >= 0 && (qsel != TB_Q_C || n != c->conc_sel)) continue;
        const ConcSlot &s = c->cs[n];
        const double *pin = c->obc.inflow ? (const double *)s.cin : nullptr;
        const double *pout = c->obc.outflow ? (const double *)s.cout : nullptr;
        const double Cx = -1.0 / (s.p.pe * c->dEta * c->dEta);
        hipLaunchKernelGGL(k_open_rhs, plane_grid(c), dim3(256), 0, c->s, c->d, c->g, (int)TB_Q_C,
                           (const uint8_t *)c->st[3], s.rhs, pin, pout, Cx);
    }
    TB_LAUNCH_CHECK(c);
    return 0;
}

int tb_open_store_old(tb_ctx *c) {
    if (!c->obc.outflow || !c->outp[0]) return 0;
    const long n = plane_n(c);
    for (int q = 0; q < 3; q++)
        hipLaunchKernelGGL(k_plane_copy, plane_grid(c), dim3(256), 0, c->s, c->outp_old[q],
                           (const double *)c->outp[q], n);
    for (int m = 0; m < c->nconc; m++)
        hipLaunchKernelGGL(k_plane_copy, plane_grid(c), dim3(256), 0, c->s, c->cs[m].cout_old,
                           (const double *)c->cs[m].cout, n);
    TB_LAUNCH_CHECK(c);
    return 0;
}

int tb_open_rk_average(tb_ctx *c, double old_frac, double new_frac) {
    if (!c->obc.outflow || !c->outp[0]) return 0;
    const long n = plane_n(c);
    for (int q = 0; q < 3; q++)
        hipLaunchKernelGGL(k_plane_rk, plane_grid(c), dim3(256), 0, c->s, c->outp[q],
                           (const double *)c->outp_old[q], old_frac, new_frac, n);
    for (int m = 0; m < c->nconc; m++)
        hipLaunchKernelGGL(k_plane_rk, plane_grid(c), dim3(256), 0, c->s, c->cs[m].cout,
                           (const double *)c->cs[m].cout_old, old_frac, new_frac, n);
    TB_LAUNCH_CHECK(c);
    return 0;
}
