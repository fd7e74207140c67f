// tb_ctx.hip -- context lifecycle, grid arrays, field transfers, reductions
// and the vector sweeps of the RK step (store_old / rk_average).
#include <chrono>
#include <cmath>
#include <cstring>

#include "tb_common.h"

// ---------------------------------------------------------------------------
// Grid_generate_mesh_coordinates (Grid.c:470-591, coordinates accumulated by
// repeated addition) and Grid_compute_metric_coefficients (Grid.c:595-708).
static void build_grid(tb_ctx *c) {
    int NX = c->d.NX, NY = c->d.NY, NZ = c->d.NZ;
    int len[21] = {NX, NY, NZ, NX, NY, NZ, NX, NY, NZ, NX, NY, NZ, NX, NY, NZ, NX, NY, NZ, NX, NY, NZ};
    for (int a = 0; a < 21; a++) c->hgrid[a].assign(len[a], 0.0);
    auto &xc = c->hgrid[0], &yc = c->hgrid[1], &zc = c->hgrid[2];
    auto &xu = c->hgrid[3], &yu = c->hgrid[4], &zu = c->hgrid[5];
    auto &xv = c->hgrid[6], &yv = c->hgrid[7], &zv = c->hgrid[8];
    auto &xw = c->hgrid[9], &yw = c->hgrid[10], &zw = c->hgrid[11];
    auto &dxc = c->hgrid[12], &dyc = c->hgrid[13], &dzc = c->hgrid[14];
    auto &mxc = c->hgrid[15], &myc = c->hgrid[16], &mzc = c->hgrid[17];
    auto &mxu = c->hgrid[18], &myv = c->hgrid[19], &mzw = c->hgrid[20];
    c->dEta = 1.0 / (double)c->ncx;
    c->dXi = 1.0 / (double)c->ncy;
    c->dPhi = 1.0 / (double)c->ncz;
    c->dx = c->cfg.lx * c->dEta;
    c->dy = c->cfg.ly * c->dXi;
    c->dz = c->cfg.lz * c->dPhi;
    const double dx = c->dx, dy = c->dy, dz = c->dz;
    auto fill = [](std::vector<double> &a, double start, double step) {
        double g = start;
        for (size_t i = 0; i < a.size(); i++) { a[i] = g; g += step; }
    };
    fill(xu, 0.0, dx); fill(yu, 0.5 * dy, dy); fill(zu, 0.5 * dz, dz);
    fill(xv, 0.5 * dx, dx); fill(yv, 0.0, dy); fill(zv, 0.5 * dz, dz);
    fill(xw, 0.5 * dx, dx); fill(yw, 0.5 * dy, dy); fill(zw, 0.0, dz);
    fill(xc, 0.5 * dx, dx); fill(yc, 0.5 * dy, dy); fill(zc, 0.5 * dz, dz);
    const double dEta = c->dEta, dXi = c->dXi, dPhi = c->dPhi;
    for (int i = 1; i < NX - 1; i++) { mxu[i] = dEta / (xc[i] - xc[i - 1]); dxc[i] = xu[i + 1] - xu[i]; }
    dxc[0] = xu[1] - xu[0];
    dxc[NX - 1] = xu[NX - 1] - xu[NX - 2];
    mxu[0] = 0.5 * dEta / (xc[0] - xu[0]);
    mxu[NX - 1] = 0.5 * dEta / (xu[NX - 1] - xc[NX - 2]);
    for (int j = 1; j < NY - 1; j++) { myv[j] = dXi / (yc[j] - yc[j - 1]); dyc[j] = yv[j + 1] - yv[j]; }
    dyc[0] = yv[1] - yv[0];
    dyc[NY - 1] = yv[NY - 1] - yv[NY - 2];
    myv[0] = 0.5 * dXi / (yc[0] - yv[0]);
    myv[NY - 1] = 0.5 * dXi / (yv[NY - 1] - yc[NY - 2]);
    for (int k = 1; k < NZ - 1; k++) { mzw[k] = dPhi / (zc[k] - zc[k - 1]); dzc[k] = zw[k + 1] - zw[k]; }
    dzc[0] = zw[1] - zw[0];
    dzc[NZ - 1] = zw[NZ - 1] - zw[NZ - 2];
    mzw[0] = 0.5 * dPhi / (zc[0] - zw[0]);
    mzw[NZ - 1] = 0.5 * dPhi / (zw[NZ - 1] - zc[NZ - 2]);
    for (int i = 0; i < NX - 1; i++) mxc[i] = dEta / (xu[i + 1] - xu[i]);
    mxc[NX - 1] = dEta / (xc[NX - 1] - xc[NX - 2]);
    for (int j = 0; j < NY - 1; j++) myc[j] = dXi / (yv[j + 1] - yv[j]);
    myc[NY - 1] = dXi / (yc[NY - 1] - yc[NY - 2]);
    for (int k = 0; k < NZ - 1; k++) mzc[k] = dPhi / (zc[k + 1] - zc[k]);   // zc, as Grid.c:697
    mzc[NZ - 1] = dPhi / (zc[NZ - 1] - zc[NZ - 2]);
}

static int dev_alloc(tb_ctx *c, void **p, size_t bytes) {
    TB_CHECK(c, hipMalloc(p, bytes));
    TB_CHECK(c, hipMemsetAsync(*p, 0, bytes, c->s));
    c->bytes += (int64_t)bytes;
    return 0;
}

int tb_helm_inc_alloc(tb_ctx *c, int q) {
    HelmHistory &h = c->helm[q];
    if (h.buf && h.inc) return 0;
    h.inc_ok = false;
    if (!h.buf && dev_alloc(c, (void **)&h.buf, sizeof(double) * (size_t)c->d.ntot)) return -1;
    if (!h.inc && dev_alloc(c, (void **)&h.inc, sizeof(float) * (size_t)c->d.ntot)) return -1;
    return 0;
}

double tb_elapsed_ms(tb_ctx *c) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
        // the event may not yet be marked complete after a spin wait
        hipEventSynchronize(c->ev1);
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
    }
    return (double)ms;
}

extern "C" tb_ctx *tb_create(const tb_config *cfg) {
    if (!cfg || cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4 || cfg->eno_order < 1 || cfg->eno_order > 3 ||
        (cfg->reinit_guard != TB_REINIT_SCALED && cfg->reinit_guard != TB_REINIT_REFERENCE))
        return nullptr;
    tb_ctx *c = new tb_ctx();
    c->cfg = *cfg;
    if (c->cfg.helm_maxit <= 0) c->cfg.helm_maxit = 200;
    if (c->cfg.pois_maxit <= 0) c->cfg.pois_maxit = 200;
    if (c->cfg.mg_pre <= 0) c->cfg.mg_pre = 2;
    if (c->cfg.mg_post <= 0) c->cfg.mg_post = 2;
    c->ncx = cfg->nx; c->ncy = cfg->ny; c->ncz = cfg->nz;
    c->d.NX = cfg->nx + 1; c->d.NY = cfg->ny + 1; c->d.NZ = cfg->nz + 1;
    c->d.px = (c->d.NX + 15) / 16 * 16;
    c->d.plane = (long)c->d.px * c->d.NY;
    c->d.ntot = c->d.plane * c->d.NZ;
    if (hipStreamCreate(&c->s) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess || hipEventCreate(&c->step_ev[0]) != hipSuccess ||
        hipEventCreate(&c->step_ev[1]) != hipSuccess) {
        delete c;
        return nullptr;
    }
    build_grid(c);
    int maxn = c->d.NX > c->d.NY ? c->d.NX : c->d.NY;
    if (c->d.NZ > maxn) maxn = c->d.NZ;
    bool ok = dev_alloc(c, (void **)&c->dgrid, sizeof(double) * 21 * (size_t)maxn) == 0;
    if (ok) {
        const double **gp[21] = {&c->g.xc, &c->g.yc, &c->g.zc, &c->g.xu, &c->g.yu, &c->g.zu,
                                 &c->g.xv, &c->g.yv, &c->g.zv, &c->g.xw, &c->g.yw, &c->g.zw,
                                 &c->g.dxc, &c->g.dyc, &c->g.dzc, &c->g.mxc, &c->g.myc, &c->g.mzc,
                                 &c->g.mxu, &c->g.myv, &c->g.mzw};
        for (int a = 0; a < 21; a++) {
            double *dst = c->dgrid + (size_t)a * maxn;
            *gp[a] = dst;
            ok = ok && hipMemcpyAsync(dst, c->hgrid[a].data(), sizeof(double) * c->hgrid[a].size(),
                                      hipMemcpyHostToDevice, c->s) == hipSuccess;
        }
    }
    size_t fb = sizeof(double) * (size_t)c->d.ntot;
    for (int f = 0; ok && f < TB_F_COUNT; f++) ok = dev_alloc(c, (void **)&c->f[f], fb) == 0;
    for (int q = 0; ok && q < 4; q++) ok = dev_alloc(c, (void **)&c->old[q], fb) == 0;
    for (int q = 0; ok && q < 4; q++) ok = dev_alloc(c, (void **)&c->st[q], (size_t)c->d.ntot) == 0;
    for (int a = 0; ok && a < 8; a++) ok = dev_alloc(c, (void **)&c->kv[a], fb) == 0;
    for (int a = 0; ok && a < GUESS_HIST; a++) ok = dev_alloc(c, (void **)&c->phi_hist[a], fb) == 0;
    for (int a = 0; ok && a < GUESS_HIST; a++)
        ok = dev_alloc(c, (void **)&c->hist32[a], sizeof(float) * (size_t)c->d.ntot) == 0;
    ok = ok && dev_alloc(c, (void **)&c->red, sizeof(double) * RED_SLOTS * RED_MAX_BLOCKS) == 0;
    ok = ok && dev_alloc(c, (void **)&c->scal, sizeof(double) * S_COUNT) == 0;
    ok = ok && hipHostMalloc((void **)&c->hscal, sizeof(double) * S_COUNT, hipHostMallocCoherent) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&c->hseq, sizeof(unsigned long long), hipHostMallocCoherent) == hipSuccess;
    if (ok) *c->hseq = 0;
    if (ok) {
        ConcSlot &s0 = c->cs[0];
        s0.c = c->f[TB_F_C]; s0.conv = c->f[TB_F_CONV_C]; s0.rhs = c->f[TB_F_RHS_C]; s0.old = c->old[3];
        s0.p.type = c->cfg.conc_type;
        s0.p.pe = c->cfg.pe;
        s0.p.v_settle = c->cfg.v_settle;
        s0.p.alpha = 1.0;
        s0.p.phi_max = 0.0;
    }
    ok = ok && hipStreamSynchronize(c->s) == hipSuccess;
    if (!ok) {
        tb_destroy(c);
        return nullptr;
    }
    if (tb_setup_geometry(c, nullptr) != 0) {
        tb_destroy(c);
        return nullptr;
    }
    return c;
}

static void free_ib(tb_ctx *c, int q) {
    hipFree(c->ib[q]); hipFree(c->ib_row[q]); hipFree(c->ib_cnt[q]); hipFree(c->ib_col[q]); hipFree(c->ib_val[q]);
    c->ib[q] = nullptr; c->ib_row[q] = nullptr; c->ib_cnt[q] = nullptr; c->ib_col[q] = nullptr; c->ib_val[q] = nullptr;
    c->nib[q] = 0;
}

extern "C" void tb_destroy(tb_ctx *c) {
    if (!c) return;
    if (c->s) hipStreamSynchronize(c->s);
    tb_mg_free(c);
    hipFree(c->dgrid);
    hipFree(c->eno_w);
    if (c->cs[0].c) {   // field 0's buffers back in place; the others freed here
        tb_conc_sync(c);
        c->f[TB_F_C] = c->cs[0].c; c->f[TB_F_CONV_C] = c->cs[0].conv;
        c->f[TB_F_RHS_C] = c->cs[0].rhs; c->old[3] = c->cs[0].old;
    }
    for (int n = 1; n < TB_NCONC_MAX; n++) {
        hipFree(c->cs[n].c); hipFree(c->cs[n].conv); hipFree(c->cs[n].rhs); hipFree(c->cs[n].old);
    }
    tb_open_free(c);
    for (int q = 0; q < 3; q++) { hipFree(c->helm[q].buf); hipFree(c->helm[q].inc); }
    // slot conc_sel's buffers are helm[3]'s (synced)
    for (int n = 0; n < TB_NCONC_MAX; n++) { hipFree(c->cs[n].h.buf); hipFree(c->cs[n].h.inc); }
    hipFree(c->ctot);
    tb_free_deposits(c, 0);
    for (int f = 0; f < TB_F_COUNT; f++) hipFree(c->f[f]);
    for (int q = 0; q < 4; q++) { hipFree(c->old[q]); hipFree(c->st[q]); free_ib(c, q); }
    for (int a = 0; a < 8; a++) hipFree(c->kv[a]);
    for (int a = 0; a < GUESS_HIST; a++) { hipFree(c->phi_hist[a]); hipFree(c->hist32[a]); }
    for (auto &e : c->prof_ev) hipEventDestroy(e);
    hipFree(c->red);
    hipFree(c->scal);
    if (c->hscal) hipHostFree(c->hscal);
    if (c->hseq) hipHostFree(c->hseq);
    for (int e = 0; e < 2; e++)
        if (c->net_ev[e]) hipEventDestroy(c->net_ev[e]);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (int e = 0; e < 2; e++)
        if (c->step_ev[e]) hipEventDestroy(c->step_ev[e]);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
}

extern "C" const char *tb_last_error(tb_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int tb_device_bytes(tb_ctx *c, int64_t *bytes) {
    if (!c || !bytes) return -1;
    *bytes = c->bytes;
    return 0;
}

void tb_free_ib(tb_ctx *c, int q) { free_ib(c, q); }

// The host copy of one 1-D grid array (the order of tb_grid_array's which).
extern "C" int tb_grid_array(tb_ctx *c, int which, double *out, int capacity) {
    if (!c || which < 0 || which >= 21 || !out) {
        if (c) c->err = "tb_grid_array: bad argument";
        return -1;
    }
    const std::vector<double> &a = c->hgrid[which];
    if ((size_t)capacity < a.size()) {
        c->err = "tb_grid_array: capacity too small";
        return -1;
    }
    memcpy(out, a.data(), sizeof(double) * a.size());
    return (int)a.size();
}

// ---------------------------------------------------------------------------
extern "C" int tb_upload_field(tb_ctx *c, int f, const double *host) {
    if (!c || f < 0 || f >= TB_F_COUNT || !host) { if (c) c->err = "tb_upload_field: bad argument"; return -1; }
    const Dims &d = c->d;
    TB_CHECK(c, hipMemcpy2DAsync(c->f[f], sizeof(double) * d.px, host, sizeof(double) * d.NX,
                                 sizeof(double) * d.NX, (size_t)d.NY * d.NZ, hipMemcpyHostToDevice, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    if (f == TB_F_PHI) c->n_hist = c->gram_n = c->guess_nr = 0;
    tb_field_written(c, f);
    return 0;
}

extern "C" int tb_download_field(tb_ctx *c, int f, double *host) {
    if (!c || f < 0 || f >= TB_F_COUNT || !host) { if (c) c->err = "tb_download_field: bad argument"; return -1; }
    const Dims &d = c->d;
    TB_CHECK(c, hipMemcpy2DAsync(host, sizeof(double) * d.NX, c->f[f], sizeof(double) * d.px,
                                 sizeof(double) * d.NX, (size_t)d.NY * d.NZ, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_download_flags(tb_ctx *c, int q, uint8_t *out) {
    if (!c || q < 0 || q > 3 || !out) { if (c) c->err = "tb_download_flags: bad argument"; return -1; }
    const Dims &d = c->d;
    TB_CHECK(c, hipMemcpy2DAsync(out, d.NX, c->st[q], d.px, d.NX, (size_t)d.NY * d.NZ, hipMemcpyDeviceToHost, c->s));
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_ib_count(tb_ctx *c, int q) {
    if (!c || q < 0 || q > 3) return -1;
    return c->nib[q];
}

extern "C" int tb_download_ib_table(tb_ctx *c, int q, tb_ib_node *out, int capacity) {
    if (!c || q < 0 || q > 3 || (!out && capacity > 0)) { if (c) c->err = "tb_download_ib_table: bad argument"; return -1; }
    int n = capacity < c->nib[q] ? capacity : c->nib[q];
    if (n > 0) {
        TB_CHECK(c, hipMemcpyAsync(out, c->ib[q], sizeof(tb_ib_node) * (size_t)n, hipMemcpyDeviceToHost, c->s));
        TB_CHECK(c, hipStreamSynchronize(c->s));
    }
    return n;
}

extern "C" int tb_reinit_iterations(tb_ctx *c) { return c ? c->reinit_its : -1; }

// ---------------------------------------------------------------------------
// live per-kernel timing: a pool of event pairs recorded around selected
// launches on the context's stream, resolved when the pool fills or on read.
constexpr int PROF_CAP = 2048;

void tb_prof_begin(tb_ctx *c, int kid) {
    if (c->prof_ev.empty()) {
        c->prof_ev.resize(2 * PROF_CAP);
        c->prof_kid.resize(PROF_CAP);
        for (auto &e : c->prof_ev) hipEventCreate(&e);
    }
    if (c->prof_n >= PROF_CAP) tb_prof_flush(c);
    hipEventRecord(c->prof_ev[2 * c->prof_n], c->s);
    c->prof_kid[c->prof_n] = kid;
}

void tb_prof_end(tb_ctx *c, int kid) {
    (void)kid;
    hipEventRecord(c->prof_ev[2 * c->prof_n + 1], c->s);
    c->prof_n++;
}

int tb_prof_flush(tb_ctx *c) {
    if (c->prof_n == 0) return 0;
    TB_CHECK(c, hipEventSynchronize(c->prof_ev[2 * (c->prof_n - 1) + 1]));
    for (int a = 0; a < c->prof_n; a++) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->prof_ev[2 * a], c->prof_ev[2 * a + 1]);
        int k = c->prof_kid[a];
        if (k >= 0 && k < TB_K_COUNT) { c->prof_ms[k] += ms; c->prof_calls[k]++; }
    }
    c->prof_n = 0;
    return 0;
}

extern "C" int tb_profile_enable(tb_ctx *c, int on) {
    if (!c) return -1;
    if (tb_prof_flush(c)) return -1;
    c->prof_on = on != 0;
    for (int k = 0; k < TB_K_COUNT; k++) { c->prof_ms[k] = 0.0; c->prof_calls[k] = 0; }
    return 0;
}

extern "C" int tb_profile_read(tb_ctx *c, int kernel, double *total_ms, int64_t *calls) {
    if (!c || kernel < 0 || kernel >= TB_K_COUNT) return -1;
    if (tb_prof_flush(c)) return -1;
    if (total_ms) *total_ms = c->prof_ms[kernel];
    if (calls) *calls = c->prof_calls[kernel];
    return 0;
}

extern "C" int tb_set_poisson_mode(tb_ctx *c, int mode, double rtol) {
    if (!c || mode < TB_POIS_FP64_PCG || mode > TB_POIS_LEARNED_IR || !(rtol > 0.0)) {
        if (c) c->err = "tb_set_poisson_mode: bad argument";
        return -1;
    }
    c->cfg.pois_mode = mode;
    c->cfg.pois_rtol = rtol;
    return 0;
}

extern "C" int tb_set_helmholtz_method(tb_ctx *c, int method) {
    if (!c || (method != TB_HELM_GS && method != TB_HELM_BICGSTAB)) {
        if (c) c->err = "tb_set_helmholtz_method: bad method";
        return -1;
    }
    c->helm_method = method;
    return 0;
}

extern "C" int tb_set_helmholtz_rtol(tb_ctx *c, double rtol) {
    if (!c || !(rtol > 0.0)) { if (c) c->err = "tb_set_helmholtz_rtol: bad argument"; return -1; }
    c->cfg.helm_rtol = rtol;
    return 0;
}

// ---------------------------------------------------------------------------
// reductions

__global__ void k_reduce_finish(const double *__restrict__ part, double *__restrict__ out, int nblocks,
                                int nslots, int slot0) {
    __shared__ double sh[RED_SLOTS * 16];
    double v[RED_SLOTS];
#pragma unroll
    for (int a = 0; a < RED_SLOTS; a++) {
        double s = 0.0;
        if (a < nslots)
            for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[a * RED_MAX_BLOCKS + b];
        v[a] = s;
    }
    block_sum<RED_SLOTS>(v, sh);
    if (threadIdx.x == 0)
        for (int a = 0; a < nslots; a++) out[slot0 + a] = v[a];
}

// Scalar read-back without a blocking stream synchronisation: one thread
// copies scal[first, first + count) into the pinned mirror, fences at
// system scope and then writes the sequence number into a pinned word, on
// which the host spins (a blocking wait wakes the host tens of microseconds
// after the GPU is done, and the GPU idles that long at every convergence
// test).  Every kernel queued before it has finished once the word changes.
// A queue that stops (a failed kernel) never writes it: after one second of
// spinning the host falls back to hipStreamSynchronize, which reports the
// error.
__global__ void k_publish(const double *__restrict__ scal, int first, int count, double *hscal,
                          unsigned long long *hseq, unsigned long long seq) {
    if (threadIdx.x != 0) return;
    for (int a = 0; a < count; a++) hscal[first + a] = scal[first + a];
    __threadfence_system();
    *(volatile unsigned long long *)hseq = seq;
}

int tb_read_scalars(tb_ctx *c, int first, int count) {
    const unsigned long long want = ++c->pub_seq;
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, c->s, (const double *)c->scal, first, count, c->hscal, c->hseq,
                       want);
    TB_LAUNCH_CHECK(c);
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned n = 1; __atomic_load_n(c->hseq, __ATOMIC_ACQUIRE) != want; n++) {
        if ((n & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
            TB_CHECK(c, hipStreamSynchronize(c->s));
            if (__atomic_load_n(c->hseq, __ATOMIC_ACQUIRE) != want) {
                c->err = "scalar read-back: the device did not publish";
                return -1;
            }
            break;
        }
        __builtin_ia32_pause();
    }
    return 0;
}

// wait until the context's stream has drained (spin, as tb_read_scalars)
int tb_sync(tb_ctx *c) { return tb_read_scalars(c, 0, 0); }

int tb_reduce_finish(tb_ctx *c, int nblocks, int nslots, int slot0) {
    hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(1024), 0, c->s, c->red, c->scal, nblocks, nslots, slot0);
    TB_LAUNCH_CHECK(c);
    return 0;
}

constexpr int DOT_BLOCKS = 1024;

__global__ void k_dot(const double *__restrict__ a, const double *__restrict__ b, long n, double *part) {
    __shared__ double sh[16];
    double v[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[0] += a[i] * b[i];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

int tb_dot(tb_ctx *c, const double *a, const double *b, int slot) {
    hipLaunchKernelGGL(k_dot, dim3(DOT_BLOCKS), dim3(256), 0, c->s, a, b, c->d.ntot, c->red);
    TB_LAUNCH_CHECK(c);
    return tb_reduce_finish(c, DOT_BLOCKS, 1, slot);
}

// ---------------------------------------------------------------------------
// store_old (VecCopy, Velocity.c:297-302 / Conc.c:997-1002) and the RK
// average y = a*x + b*y (VecAXPBY, Velocity.c:2047-2055 / Conc.c:1007-1017).

__global__ void k_axpby4(double *y0, double *y1, double *y2, double *y3, const double *x0, const double *x1,
                         const double *x2, const double *x3, double a, double b, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    y0[i] = a * x0[i] + b * y0[i];
    y1[i] = a * x1[i] + b * y1[i];
    y2[i] = a * x2[i] + b * y2[i];
    y3[i] = a * x3[i] + b * y3[i];
}

__global__ void k_axpby1(double *y, const double *x, double a, double b, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = a * x[i] + b * y[i];
}

extern "C" int tb_store_old(tb_ctx *c) {
    if (!c) return -1;
    const size_t bytes = sizeof(double) * c->d.ntot;
    tb_conc_sync(c);
    for (int q = 0; q < 3; q++)
        TB_CHECK(c, hipMemcpyAsync(c->old[q], c->f[TB_F_U + q], bytes, hipMemcpyDeviceToDevice, c->s));
    for (int n = 0; n < c->nconc; n++)
        TB_CHECK(c, hipMemcpyAsync(c->cs[n].old, c->cs[n].c, bytes, hipMemcpyDeviceToDevice, c->s));
    return tb_open_store_old(c);
}

extern "C" int tb_rk_average(tb_ctx *c, double old_frac, double new_frac) {
    if (!c) return -1;
    long n = c->d.ntot;
    const dim3 grid((unsigned)((n + 255) / 256));
    tb_conc_sync(c);
    hipLaunchKernelGGL(k_axpby4, grid, dim3(256), 0, c->s, c->f[TB_F_U], c->f[TB_F_V], c->f[TB_F_W], c->cs[0].c,
                       c->old[0], c->old[1], c->old[2], c->cs[0].old, old_frac, new_frac, n);
    for (int m = 1; m < c->nconc; m++)
        hipLaunchKernelGGL(k_axpby1, grid, dim3(256), 0, c->s, c->cs[m].c, (const double *)c->cs[m].old, old_frac,
                           new_frac, n);
    TB_LAUNCH_CHECK(c);
    return tb_open_rk_average(c, old_frac, new_frac);
}

// ---------------------------------------------------------------------------
// concentration fields

// Points the selected-field views (f[TB_F_C], f[TB_F_CONV_C], f[TB_F_RHS_C],
// old[3], cfg.pe / conc_type / v_settle and the Gauss-Seidel history of q = 3)
// at field n; the IMMERSED rows of the concentration grid are scaled by
// ap_solid, which depends on Pe, so they are rebuilt when Pe changes.
int tb_conc_use(tb_ctx *c, int n) {
    tb_conc_sync(c);   // the selected field's views and Gauss-Seidel history back into its slot
    const ConcSlot &s = c->cs[n];
    const bool pe_changed = c->cfg.pe != s.p.pe;
    c->f[TB_F_C] = s.c; c->f[TB_F_CONV_C] = s.conv; c->f[TB_F_RHS_C] = s.rhs; c->old[TB_Q_C] = s.old;
    c->cfg.pe = s.p.pe;
    c->cfg.conc_type = s.p.type;
    c->cfg.v_settle = s.p.v_settle;
    c->helm[TB_Q_C] = s.h;
    c->conc_sel = n;
    if (pe_changed && c->nib[TB_Q_C] > 0) return tb_ib_rows(c, TB_Q_C);
    return 0;
}

// frees a field-sized buffer of T and clears its pointer
template <typename T>
static void dev_release(tb_ctx *c, T *&p) {
    if (!p) return;
    hipFree(p);
    p = nullptr;
    c->bytes -= (int64_t)(sizeof(T) * (size_t)c->d.ntot);
}

extern "C" int tb_set_conc_fields(tb_ctx *c, int nconc, const tb_conc_field *fields, int hindered) {
    if (!c || nconc < 1 || nconc > TB_NCONC_MAX || !fields) {
        if (c) c->err = "tb_set_conc_fields: bad argument";
        return -1;
    }
    for (int n = 0; n < nconc; n++) {
        const tb_conc_field &p = fields[n];
        if ((p.type != TB_SALINITY && p.type != TB_PARTICLE) || !(p.pe > 0.0) || !std::isfinite(p.pe) ||
            !std::isfinite(p.v_settle) || !std::isfinite(p.alpha) || !std::isfinite(p.phi_max)) {
            c->err = "tb_set_conc_fields: field " + std::to_string(n) + " has a bad type or parameter";
            return -1;
        }
    }
    if (tb_conc_use(c, 0)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    const size_t fb = sizeof(double) * (size_t)c->d.ntot;
    for (int n = 1; n < TB_NCONC_MAX; n++) {
        ConcSlot &s = c->cs[n];
        if (n < nconc) {
            if (!s.c) {
                if (dev_alloc(c, (void **)&s.c, fb) || dev_alloc(c, (void **)&s.conv, fb) ||
                    dev_alloc(c, (void **)&s.rhs, fb) || dev_alloc(c, (void **)&s.old, fb))
                    return -1;
            }
        } else {
            dev_release(c, s.c); dev_release(c, s.conv); dev_release(c, s.rhs); dev_release(c, s.old);
            dev_release(c, s.h.buf);
            dev_release(c, s.h.inc);
        }
    }
    tb_free_deposits(c, nconc);
    if (nconc > 1 && !c->ctot) {
        if (dev_alloc(c, (void **)&c->ctot, fb)) return -1;
    } else if (nconc == 1) {
        dev_release(c, c->ctot);
    }
    for (int n = 0; n < nconc; n++) {
        c->cs[n].p = fields[n];
        double *buf = c->cs[n].h.buf;   // kept, the increment's content no longer a guess
        float *inc = c->cs[n].h.inc;
        c->cs[n].h = HelmHistory();
        c->cs[n].h.buf = buf;
        c->cs[n].h.inc = inc;
    }
    c->helm[TB_Q_C] = c->cs[0].h;
    c->nconc = nconc;
    c->hindered = hindered != 0;
    if (tb_conc_use(c, 0)) return -1;   // field 0's (possibly new) parameters
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tb_select_conc(tb_ctx *c, int n) {
    if (!c || n < 0 || n >= c->nconc) {
        if (c) c->err = "tb_select_conc: no such concentration field";
        return -1;
    }
    return tb_conc_use(c, n);
}

// c_total = alpha_0 c_0, then c_total += alpha_i c_i (VecCopy, VecScale and
// VecAXPY in Conc_compute_total_concentration, Conc.c:1291-1314)
__global__ void k_total_conc(double *t, const double *c0, const double *c1, const double *c2, const double *c3,
                             double a0, double a1, double a2, double a3, int nconc, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = a0 * c0[i];
    if (nconc > 1) s = s + a1 * c1[i];
    if (nconc > 2) s = s + a2 * c2[i];
    if (nconc > 3) s = s + a3 * c3[i];
    t[i] = s;
}

extern "C" int tb_total_concentration(tb_ctx *c) {
    if (!c) return -1;
    if (c->nconc == 1) return 0;
    tb_conc_sync(c);
    const ConcSlot *s = c->cs;
    const long n = c->d.ntot;
    hipLaunchKernelGGL(k_total_conc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, c->ctot,
                       (const double *)s[0].c, (const double *)s[1].c, (const double *)s[2].c,
                       (const double *)s[3].c, s[0].p.alpha, s[1].p.alpha, s[2].p.alpha, s[3].p.alpha, c->nconc, n);
    TB_LAUNCH_CHECK(c);
    return 0;
}
