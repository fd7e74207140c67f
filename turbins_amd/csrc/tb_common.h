// tb_common.h -- shared definitions of the gfx950 library: context, field
// layout, launch geometry and small device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/turbins_gpu.h"

// ---------------------------------------------------------------------------
// Field layout in HBM: the reference's natural order (x fastest, then y, z)
// over (nx+1)(ny+1)(nz+1) points, with each x row padded to a pitch that is
// a multiple of 16 doubles so every row starts on a 128-byte boundary.
struct Dims {
    int NX, NY, NZ;     // points per axis (cells + the half cell)
    int px;             // row pitch in elements
    long plane;         // px * NY
    long ntot;          // plane * NZ
};

__host__ __device__ inline long gidx(const Dims &d, int i, int j, int k) {
    return ((long)k * d.NY + j) * d.px + i;
}

// 1-D coordinate and metric arrays (Grid.c:470-708), device pointers.
struct GridDev {
    const double *xc, *yc, *zc, *xu, *yu, *zu, *xv, *yv, *zv, *xw, *yw, *zw;
    const double *dxc, *dyc, *dzc;
    const double *mxc, *myc, *mzc, *mxu, *myv, *mzw;
};

// MyMath_interpolate_quantity, two points (MyMath.c:24-41): linear in x,
// the first value where the two abscissae are closer than 1e-8
__host__ __device__ inline double interp2d(double q0, double x0, double q1, double x1, double xi) {
    double m = (q1 - q0) / (x1 - x0);
    if (fabs(x1 - x0) < 1.0e-8) return q0;
    return m * (xi - x0) + q0;
}

// ---------------------------------------------------------------------------
// Launch geometry for point-wise sweeps: tiles of 64 x 4 points in one k
// plane.  Workgroups are handed to the 8 XCDs round robin (block b runs on
// XCD b % 8), so the linear block id is remapped to give each XCD one
// contiguous slab of tiles: neighbouring rows and planes of a stencil then
// hit the same XCD's L2.
constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;
constexpr int NXCD = 8;

struct Tiles {
    int nx, ny, nz;     // extent of the swept box (points)
    int tx, ty;         // tiles per axis
    long n;             // tiles in total
    long per_xcd;       // ceil(n / 8)
};

inline Tiles make_tiles(int nx, int ny, int nz) {
    Tiles t;
    t.nx = nx; t.ny = ny; t.nz = nz;
    t.tx = (nx + TILE_X - 1) / TILE_X;
    t.ty = (ny + TILE_Y - 1) / TILE_Y;
    t.n = (long)t.tx * t.ty * nz;
    t.per_xcd = (t.n + NXCD - 1) / NXCD;
    return t;
}

inline dim3 tiles_grid(const Tiles &t) { return dim3((unsigned)(t.per_xcd * NXCD)); }
inline dim3 tiles_block() { return dim3(TILE_X, TILE_Y); }

// Returns false for threads outside the box (or remapped past the last tile).
__device__ inline bool tile_point(const Tiles &t, int &i, int &j, int &k) {
    long b = blockIdx.x;
    long lt = (b % NXCD) * t.per_xcd + b / NXCD;
    if (lt >= t.n) return false;
    int tx = (int)(lt % t.tx);
    long r = lt / t.tx;
    int ty = (int)(r % t.ty);
    k = (int)(r / t.ty);
    i = tx * TILE_X + threadIdx.x;
    j = ty * TILE_Y + threadIdx.y;
    return i < t.nx && j < t.ny;
}

// The same for virtual block vb of a loop over the tiles (vb < per_xcd * 8).
__device__ inline bool tile_point_at(const Tiles &t, long vb, int &i, int &j, int &k) {
    long lt = (vb % NXCD) * t.per_xcd + vb / NXCD;
    if (lt >= t.n) return false;
    int tx = (int)(lt % t.tx);
    long r = lt / t.tx;
    int ty = (int)(r % t.ty);
    k = (int)(r / t.ty);
    i = tx * TILE_X + threadIdx.x;
    j = ty * TILE_Y + threadIdx.y;
    return i < t.nx && j < t.ny;
}

// Same tiles, numbered with k fastest: the blocks an XCD runs back to back
// cover one 64 x 4 column plane after plane, so a stencil reaching several
// planes in z finds them in that XCD's L2 (the x-fastest order above revisits
// a column only after a whole plane of tiles).
__device__ inline bool tile_point_kfast(const Tiles &t, int &i, int &j, int &k) {
    long b = blockIdx.x;
    long lt = (b % NXCD) * t.per_xcd + b / NXCD;
    if (lt >= t.n) return false;
    k = (int)(lt % t.nz);
    long r = lt / t.nz;
    int tx = (int)(r % t.tx);
    int ty = (int)(r / t.tx);
    i = tx * TILE_X + threadIdx.x;
    j = ty * TILE_Y + threadIdx.y;
    return i < t.nx && j < t.ny;
}

// Column tiles for k-marching stencils: a 64 x 4 block owns a tile of
// columns and walks KCHUNK planes upwards, keeping the k-1, k, k+1 values of
// its column in registers, so each plane is read from HBM once.  Only full
// tiles are used; the columns they leave (a point box of 513 x 257 leaves
// column 512 and row 256) are listed 256 to a block, and the last k chunk
// takes up to KCHUNK/2 extra planes (129 planes: 8 chunks, not 9).  Each
// block is a chain of plane steps, so nearly empty tiles would cost as much
// time as full ones.
constexpr int KCHUNK = 16;

struct KTiles {
    int nx, ny, nz;
    int tx, ty, kc;     // full tiles per axis, k chunks
    int fx, fy;         // extent of the full tiles
    int ne, eb;         // columns outside the full tiles, blocks listing them (per chunk)
    long n, per_xcd;    // blocks: (tx * ty + eb) * kc
};

inline KTiles make_ktiles(int nx, int ny, int nz) {
    KTiles t;
    t.nx = nx; t.ny = ny; t.nz = nz;
    t.tx = nx / TILE_X;
    t.ty = ny / TILE_Y;
    t.fx = t.tx * TILE_X;
    t.fy = t.ty * TILE_Y;
    t.ne = nx * ny - t.fx * t.fy;
    t.eb = (t.ne + TILE_X * TILE_Y - 1) / (TILE_X * TILE_Y);
    t.kc = (nz + KCHUNK / 2) / KCHUNK;
    if (t.kc < 1) t.kc = 1;
    t.n = (long)(t.tx * t.ty + t.eb) * t.kc;
    t.per_xcd = (t.n + NXCD - 1) / NXCD;
    return t;
}

inline dim3 ktiles_grid(const KTiles &t) { return dim3((unsigned)(t.per_xcd * NXCD)); }

// Column (i, j) and plane range [k0, k1) of this thread.  Threads with
// i >= nx or j >= ny, and whole blocks remapped past the last tile, get
// active = false (they must still join block reductions).
__device__ inline void ktile_column(const KTiles &t, int &i, int &j, int &k0, int &k1, bool &active) {
    int b = blockIdx.x;
    int per = (int)t.per_xcd;
    int lt = (b % NXCD) * per + b / NXCD;
    i = j = k0 = k1 = 0;
    active = false;
    if (lt >= t.n) return;
    const int nmain = t.tx * t.ty;
    const int per_chunk = nmain + t.eb;
    const int kc = lt / per_chunk;
    // the listed (uncoalesced, slower) blocks of a chunk come first, so they
    // do not form the tail of the launch
    const int r0 = lt - kc * per_chunk;
    const int r = r0 < t.eb ? r0 + nmain : r0 - t.eb;
    k0 = kc * KCHUNK;
    k1 = (kc == t.kc - 1) ? t.nz : k0 + KCHUNK;
    if (r < nmain) {
        i = (r % t.tx) * TILE_X + threadIdx.x;
        j = (r / t.tx) * TILE_Y + threadIdx.y;
        active = true;
        return;
    }
    // columns outside the full tiles: i >= fx (all j) first, then j >= fy (i < fx)
    int e = (r - nmain) * (TILE_X * TILE_Y) + threadIdx.y * TILE_X + threadIdx.x;
    if (e >= t.ne) return;
    const int wx = t.nx - t.fx;
    if (e < wx * t.ny) {
        i = t.fx + e % wx;
        j = e / wx;
    } else {
        e -= wx * t.ny;
        i = e % t.fx;
        j = t.fy + e / t.fx;
    }
    active = true;
}

// ---------------------------------------------------------------------------
// Reductions: each block writes partial sums; a one-block kernel adds the
// partials in a fixed order, so results do not depend on scheduling.
// streaming store: bypasses the caches' retention so the read streams of a
// stencil keep their halo lines (7-point FP64 stencil at 513x257x129 on
// gfx950: 77 -> 64 us)
template <typename T>
__device__ inline void st_stream(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

constexpr int RED_MAX_BLOCKS = 1 << 18;   // per slot; one partial per workgroup
constexpr int RED_SLOTS = 4;        // up to 4 simultaneous sums per launch

template <typename T>
__device__ inline T warp_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sums NV values over the block; thread 0 of the block gets the totals.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *sh /* [NV*16] */) {
    int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    int wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    int nw = (blockDim.x * blockDim.y + 63) >> 6;
#pragma unroll
    for (int a = 0; a < NV; a++) v[a] = warp_sum(v[a]);
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < NV; a++) sh[a * 16 + wid] = v[a];
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int a = 0; a < NV; a++) {
            double s = (lane < nw) ? sh[a * 16 + lane] : 0.0;
            v[a] = warp_sum(s);
        }
    }
}

// ---------------------------------------------------------------------------
// GMG level description (cell-centred arrays)
// history depth of the Galerkin initial guess, and its directions with the
// network's
constexpr int GUESS_HIST = 10;     // stored solutions (upper bound of tb_set_guess_history)
constexpr int GUESS_DEFAULT = 8;
constexpr int GUESS_DIRS = GUESS_HIST + 1;

struct Level {
    int nx, ny, nz;       // cells
    int px;               // row pitch (elements)
    long plane, ntot;
    double hx, hy, hz;
    int cx, cy, cz;       // halved towards the next level
    double *x64, *b64, *r64;
    float *x32, *b32, *r32;
    double *a64;          // second iterate buffer of levels smoothed by fused pairs, else null
    float *a32;
    bool pairs;           // smoothed by fused sweep pairs
};

// ---------------------------------------------------------------------------
// One concentration field (a Concentration object, Conc.c:22-217): its data,
// convective term, right-hand side and stage-start copy, its parameters, and
// the Gauss-Seidel history of its Helmholtz solves (HelmHistory).
struct HelmHistory {
    // pair count and residual contraction per pair of the last solves, by the
    // parity of the solve count: the solves alternate between the two RK
    // stages, whose extrapolated guesses differ in quality
    int sweeps[2] = {0, 0};
    double rate[2] = {0.0, 0.0};
    unsigned nsolve = 0;
    double *buf = nullptr;       // second iterate buffer of the solves (allocated on first use, with inc)
    float *inc = nullptr;        // the last solve's increment x_new - x_start, in FP32
    bool inc_ok = false;         // inc belongs to this field's current trajectory
};

struct ConcSlot {
    double *c = nullptr, *conv = nullptr, *rhs = nullptr, *old = nullptr;
    double *dep = nullptr;       // deposited height per (i, k) column, x fastest (allocated on first use)
    tb_conc_field p{};
    HelmHistory h;
    // open x faces: inflow, outflow and stage-start outflow planes (NY*NZ, [k][j])
    double *cin = nullptr, *cout = nullptr, *cout_old = nullptr;
};

struct tb_ctx {
    tb_config cfg;
    std::string err;
    int ncx, ncy, ncz;
    Dims d;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;       // each solve's own timing
    hipEvent_t step_ev[2] = {nullptr, nullptr};    // tb_step's timing
    int64_t bytes = 0;

    // host copies of the 1-D grid arrays, and their device copies
    std::vector<double> hgrid[21];
    double *dgrid = nullptr;
    GridDev g;
    double dEta, dXi, dPhi, dx, dy, dz;

    double *f[TB_F_COUNT] = {};
    double *old[4] = {};
    uint8_t *st[4] = {};

    // concentration fields; the selected one's buffers are f[TB_F_C],
    // f[TB_F_CONV_C], f[TB_F_RHS_C] and old[3], its parameters cfg.pe,
    // cfg.conc_type and cfg.v_settle (tb_conc_use)
    ConcSlot cs[TB_NCONC_MAX];
    int nconc = 1;
    int conc_sel = 0;
    int hindered = 0;            // params->hindered_settling
    double *ctot = nullptr;      // total concentration (nconc > 1)

    // open x faces (tb_open.hip): planes of NY*NZ doubles indexed [k][j]
    tb_open_bc obc{};
    double *u_in = nullptr;                     // u->G_inflow
    double *outp[3] = {}, *outp_old[3] = {};    // u, v, w G_outflow, G_outflow_old
    double *open_scal = nullptr;                // device: [0] the influx u->W_influx

    // immersed boundary: node tables and the scaled matrix rows
    int nib[4] = {0, 0, 0, 0};
    tb_ib_node *ib[4] = {};
    long *ib_row[4] = {};
    int *ib_cnt[4] = {};
    long *ib_col[4] = {};        // 9 per row, sorted
    double *ib_val[4] = {};      // 9 per row
    int reinit_its = 0;
    bool has_geometry = false;

    // Krylov work vectors (full padded layout)
    double *kv[8] = {};
    double *eno_w = nullptr;     // interpolation weights of the momentum transposes (tb_eno.hip)
    int eno_wmax = 0;
    double *red = nullptr;       // partial sums [RED_SLOTS][RED_MAX_BLOCKS]
    double *scal = nullptr;      // device scalars
    double *hscal = nullptr;     // pinned host mirror of scal
    unsigned long long *hseq = nullptr;   // pinned word the device writes after publishing scalars
    unsigned long long pub_seq = 0;       // last sequence number asked for

    // multigrid
    int nlev = 0;
    Level lev[16];
    double *coarse_inv64 = nullptr;
    float *coarse_inv32 = nullptr;
    float *cg32[3] = {};         // p, q, e of the FP32 inner CG (level-0 layout)
    // V-cycles replayed from HIP graphs (0: FP64, 1: FP32); the FP64 graph is
    // tied to the level-0 vectors it was captured with
    bool use_graphs = true;
    hipGraphExec_t vc_exec[2] = {nullptr, nullptr};
    int helm_method = TB_HELM_GS;    // Helmholtz solver
    HelmHistory helm[4];         // Gauss-Seidel history of u, v, w and the selected concentration field
    double ir_floor = 1e-5;      // smallest residual reduction asked of one inner FP32 solve
    int ir_maxit = 40;           // inner iterations per refinement pass
    int ir_first_its = 0;        // iterations of the last mixed solve's first pass
    // V-cycles (FP32, FP64) whose last finest-level pair leaves the partial
    // sums of r . z in red[0 .. vc_dot_blocks); 0: that level has no pairs
    int vc_dot_blocks[2] = {0, 0};
    int mg_pre0[2] = {0, 0};     // sweeps on the finest level per V-cycle side (FP32, FP64);
    int mg_post0[2] = {0, 0};    // 0: cfg.mg_pre / cfg.mg_post as on the coarser levels
    bool mg_blocked = true;      // mid levels smoothed by one blocked launch per V-cycle side
    int coarse_n = 0;
    bool coarse_dense = true;    // coarsest level solved by its dense inverse (<= 1024 cells), else by sweeps
    int small_l0[2] = {0, 0};   // first level of the one-workgroup cycle: FP32, FP64
    bool mg_built = false;

    // Poisson history for the learned guess
    double *phi_hist[GUESS_HIST] = {};   // last Poisson solutions, newest first
    float *hist32[GUESS_HIST] = {};      // the same in FP32: the Galerkin guess's directions
    int n_hist = 0;
    // Galerkin matrix of the stored solutions, gram[i][j] = phi_i . A phi_j,
    // valid for the leading gram_n entries.  A phi_j = b_j holds to the solve
    // tolerance, so a new row is phi_k . b of the solve that made phi_0 (the
    // products the guess of that solve formed, guess_r[0..guess_nr)) and the
    // diagonal phi_0 . b (the final residual pass); no right-hand side is kept.
    double gram[GUESS_HIST][GUESS_HIST] = {};
    int gram_n = 0;
    double guess_r[GUESS_HIST] = {};
    int guess_nr = 0;
    bool phi_pinned = false;   // the phi field is the output of the last Poisson solve (corner exactly 0)
    bool phi_is_hist0 = false; // ... and that solve stored it as phi_hist[0]
    int guess_hist = GUESS_DEFAULT; // solutions in the Galerkin initial guess (tb_set_guess_history)
    tb_net *net = nullptr;       // attached learned guess (not owned)
    tb_net *net_record = nullptr;    // snapshot sink after every solve (not owned)
    // run-time judgement of the network: iterations its direction saves
    // (from the energy of the 2- and 3-direction guesses) against its cost
    bool net_open = true;
    int net_trials = 0;
    double net_saved = 0.0, net_cost_ms = 0.0, net_iter_ms = 0.0;
    double net_J2 = 0.0, net_J3 = 0.0;
    bool net_used = false;
    hipEvent_t net_ev[2] = {nullptr, nullptr};

    // live per-kernel timing (event pairs around selected launches)
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_kid;
    int prof_n = 0;
    double prof_ms[TB_K_COUNT] = {};
    int64_t prof_calls[TB_K_COUNT] = {};
};

void tb_prof_begin(tb_ctx *c, int kid);
void tb_prof_end(tb_ctx *c, int kid);
int tb_prof_flush(tb_ctx *c);

// wraps one launch statement with a timing event pair when profiling is on
#define TB_PROF(ctx, kid, stmt)                              \
    do {                                                     \
        if ((ctx)->prof_on) tb_prof_begin((ctx), (kid));     \
        stmt;                                                \
        if ((ctx)->prof_on) tb_prof_end((ctx), (kid));       \
    } while (0)

// device scalar slots (ctx->scal)
enum {
    S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_RZ, S_PQ, S_RR, S_TS, S_TT, S_RHV,
    S_BB, S_MEAN, S_CORNER, S_TMP0, S_TMP1, S_TMP2, S_SUB = 16,
    S_RES = 16 + GUESS_DIRS * GUESS_DIRS + GUESS_DIRS,   // Poisson residual: r.r, b.b, x.b
    S_COUNT = S_RES + 3
};

// error helpers -------------------------------------------------------------
#define TB_CHECK(ctx, call)                                                      \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);      \
            return -1;                                                           \
        }                                                                        \
    } while (0)

#define TB_LAUNCH_CHECK(ctx)                                                     \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess) {                                                  \
            (ctx)->err = std::string("launch: ") + hipGetErrorString(e_);       \
            return -1;                                                           \
        }                                                                        \
    } while (0)

// internal entry points shared between translation units
int tb_geometry_run(tb_ctx *c, const double *h_xz);
int tb_reduce_finish(tb_ctx *c, int nblocks, int nslots, int slot0);
int tb_read_scalars(tb_ctx *c, int first, int count);
int tb_sync(tb_ctx *c);
int tb_dot(tb_ctx *c, const double *a, const double *b, int slot);
int tb_mg_build(tb_ctx *c);
void tb_mg_free(tb_ctx *c);
int tb_vcycle64(tb_ctx *c, const double *r_pad, double *z_pad);
int tb_vcycle32(tb_ctx *c);
int tb_net_predict(tb_ctx *c, double *out, int correction_only);
int tb_net_record_one(tb_ctx *c);
double tb_elapsed_ms(tb_ctx *c);
int tb_conc_use(tb_ctx *c, int n);
// open x faces (tb_open.hip): the plane of quantity q (the selected field for
// c) while that face is open, else null; the right-hand-side terms of every
// quantity and field; the outflow planes' stage-start copy and RK average
const double *tb_in_plane(const tb_ctx *c, int q);
const double *tb_out_plane(const tb_ctx *c, int q);
int tb_open_rhs(tb_ctx *c, int q);   // q = -1: every quantity and field
int tb_open_store_old(tb_ctx *c);
int tb_open_rk_average(tb_ctx *c, double old_frac, double new_frac);
int tb_outflow_async(tb_ctx *c, double dt);   // tb_outflow_convective without the final sync
void tb_open_free(tb_ctx *c);
int tb_helm_inc_alloc(tb_ctx *c, int q);   // the increment buffer of quantity q (once)
void tb_free_deposits(tb_ctx *c, int first);
int tb_ib_rows(tb_ctx *c, int q);

// The selected field's views are the live pointers (a Helmholtz solve may
// hand the field a new buffer, trading it with a scratch vector): copies
// them back into its slot; every reader of the slot table calls it first.
inline void tb_conc_sync(tb_ctx *c) {
    ConcSlot &s = c->cs[c->conc_sel];
    s.c = c->f[TB_F_C];
    s.conv = c->f[TB_F_CONV_C];
    s.rhs = c->f[TB_F_RHS_C];
    s.old = c->old[TB_Q_C];
    s.h = c->helm[TB_Q_C];
}

// the source of the buoyancy term of the v right-hand side
// (Velocity.c:1798-1815): c_total with several fields, else field 0
inline const double *tb_buoyancy_source(tb_ctx *c) {
    tb_conc_sync(c);
    return c->nconc > 1 ? c->ctot : c->cs[0].c;
}

// Conc_settling_speed_function (Conc.c:815-836), the Richardson-Zaki form
// V_s0 (1 - c phi_max)^4 with the power taken as two squarings
__host__ __device__ inline double hindered_speed(double conc, double phi_max, double v_s0) {
    const double t = 1.0 - conc * phi_max;
    const double t2 = t * t;
    return v_s0 * (t2 * t2);
}

// Conc_set_particle_v_velocity (Conc.c:734-761): the v velocity that
// advects a particle field at a cell, from the fluid's cell-centre v
__host__ __device__ inline double particle_v(double vbc, int status, double conc, int hindered, double phi_max,
                                            double v_s0) {
    if (!hindered) return vbc + v_s0;
    return vbc + (status == TB_FLUID ? hindered_speed(conc, phi_max, v_s0) : 0.0);
}

// every entry point that writes a caller-chosen field reports it here
inline void tb_field_written(tb_ctx *c, int f) {
    if (f == TB_F_PHI) c->phi_pinned = c->phi_is_hist0 = false;
    if (f >= TB_F_U && f <= TB_F_C) c->helm[f - TB_F_U].inc_ok = false;   // a new trajectory
}
