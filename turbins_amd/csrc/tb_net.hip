// tb_net.hip -- learned initial guess for the pressure Poisson solve.
//
// A 3-level 3-D convolutional encoder-decoder maps (b, phi_{n-1}, phi_{n-2})
// to a correction of phi_{n-1}:  phi_guess = phi_{n-1} + s_phi * net(...),
// with b scaled by its RMS and the pressures by RMS(phi_{n-1}).
//
//   x0  = relu(conv(in, 16->16))        L0 (cells)      in: 3 of 16 channels used
//   x1  = relu(conv(pool(x0), 16->32))  L1 (cells / 2)
//   x1b = relu(conv(x1, 32->32))        L1
//   x2  = relu(conv(pool(x1b), 32->64)) L2 (cells / 4)
//   x2b = relu(conv(x2, 64->64))        L2
//   y1  = relu(conv([up(x2b), x1b], 96->32))   L1
//   y0  = relu(conv([up(y1), x0], 48->16))     L0
//   out = conv(y0, 16->1)               L0
//
// All convolutions are 3x3x3, stride 1, zero padding 1; pool is a 2x2x2
// mean, up is nearest-neighbour x2 (fused into the next convolution's
// operand gather, never materialised).  Activations are BF16 in NDHWC
// (channels fastest); weights are FP32 masters rounded to BF16 for the
// matrix cores; every product accumulates in FP32.
//
// The convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward  : out[v][n]  = sum_{t,c} W[t][c][n] * X[v + off(t)][c]
//              M = voxels, N = output channels, K = 27 * input channels;
//   backward data: the same kernel with the weights flipped in t and
//              transposed in (c, n), applied to the output gradient;
//   backward weights: dW[t*C + c][n] = sum_v X[v + off(t)][c] * dY[v][n]
//              M = (t, c), N = n, K = voxels.
// The 1-output-channel head uses plain FP32 kernels.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tb_common.h"

typedef uint16_t bf16_t;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__host__ __device__ inline float bf2f(bf16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// architecture
constexpr int NLAYER = 8;
constexpr int IN_CH = 16;      // input tensor channels (3 used)
struct LayerDef {
    int cin, cout;
};
static const LayerDef LAYERS[NLAYER] = {{16, 16}, {16, 32}, {32, 32}, {32, 64},
                                        {64, 64}, {96, 32}, {48, 16}, {16, 1}};
enum { L_IN = 0, L_E1, L_E1B, L_E2, L_E2B, L_D1, L_D0, L_OUT };

// ---------------------------------------------------------------------------
// operand gathers

struct Src {
    const bf16_t *A;    // Ca channels, at half resolution when a_up
    const bf16_t *B;    // Cb channels, full resolution (may be null, Cb = 0)
    int Ca, Cb, a_up;
};

__device__ inline bool tap_neighbour(int t, int x, int y, int z, int nx, int ny, int nz, int &xx, int &yy, int &zz) {
    xx = x + t % 3 - 1;
    yy = y + (t / 3) % 3 - 1;
    zz = z + t / 9 - 1;
    return t < 27 && xx >= 0 && yy >= 0 && zz >= 0 && xx < nx && yy < ny && zz < nz;
}

__device__ inline long src_offset(const Src &s, int c, int xx, int yy, int zz, int nx, int ny, const bf16_t *&base,
                                  int &stride) {
    if (c < s.Ca) {
        const int u = s.a_up;
        base = s.A;
        stride = s.Ca;
        return (((long)(zz >> u) * (ny >> u) + (yy >> u)) * (nx >> u) + (xx >> u)) * s.Ca + c;
    }
    base = s.B;
    stride = s.Cb;
    return (((long)zz * ny + yy) * nx + xx) * s.Cb + (c - s.Ca);
}

// 8 consecutive channels (c multiple of 8) of the tap-t neighbour, 0 outside
__device__ inline v4u gather8(const Src &s, int t, int c, int x, int y, int z, int nx, int ny, int nz) {
    v4u r = {0u, 0u, 0u, 0u};
    int xx, yy, zz;
    if (!tap_neighbour(t, x, y, z, nx, ny, nz, xx, yy, zz)) return r;
    const bf16_t *base;
    int stride;
    long o = src_offset(s, c, xx, yy, zz, nx, ny, base, stride);
    return *(const v4u *)(base + o);
}

__device__ inline bf16_t gather1(const Src &s, int t, int c, int x, int y, int z, int nx, int ny, int nz) {
    int xx, yy, zz;
    if (!tap_neighbour(t, x, y, z, nx, ny, nz, xx, yy, zz)) return 0;
    const bf16_t *base;
    int stride;
    long o = src_offset(s, c, xx, yy, zz, nx, ny, base, stride);
    return base[o];
}

__device__ inline void vox_xyz(long v, int nx, int ny, int &x, int &y, int &z) {
    x = (int)(v % nx);
    long r = v / nx;
    y = (int)(r % ny);
    z = (int)(r / ny);
}

// ---------------------------------------------------------------------------
// implicit-GEMM convolution, forward / backward data

struct ConvFwd {
    Src s;
    int nx, ny, nz;
    long nvox;
    const v4u *frag;        // [nchunks][NB][64 lanes] x 8 bf16
    int nchunks;            // ceil(27 * cin / 32)
    const float *bias;      // [cout] or null
    const bf16_t *mask;     // [nvox][cout] or null: output zeroed where mask <= 0
    int relu;
    bf16_t *out;            // [nvox][cout]
};

// 4 waves per workgroup, each wave 32 voxels (two 16-row M blocks) x all
// NB*16 output channels; K walks (tap, channel) in steps of 32.
template <int NB>
__global__ void __launch_bounds__(256) k_conv_fwd(ConvFwd a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long v0 = ((long)blockIdx.x * 4 + wave) * 32;
    const int cin = a.s.Ca + a.s.Cb;
    int vx[2], vy[2], vz[2];
    bool vok[2];
#pragma unroll
    for (int mb = 0; mb < 2; mb++) {
        long v = v0 + mb * 16 + (lane & 15);
        vok[mb] = v < a.nvox;
        vox_xyz(vok[mb] ? v : 0, a.nx, a.ny, vx[mb], vy[mb], vz[mb]);
    }
    v4f acc[2][NB];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = (v4f){0.f, 0.f, 0.f, 0.f};
    int kk = 8 * (lane >> 4);
    int t = kk / cin, c = kk - (kk / cin) * cin;
    for (int ch = 0; ch < a.nchunks; ch++) {
        v4u av[2];
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
            av[mb] = (v4u){0u, 0u, 0u, 0u};
            if (vok[mb]) av[mb] = gather8(a.s, t, c, vx[mb], vy[mb], vz[mb], a.nx, a.ny, a.nz);
        }
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const v4u bw = a.frag[((long)ch * NB + nb) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 2; mb++)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, av[mb]),
                                                                      __builtin_bit_cast(v8bf, bw), acc[mb][nb], 0, 0, 0);
        }
        c += 32;
        while (c >= cin) { c -= cin; t++; }
    }
    const int cout = NB * 16;
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const long v = v0 + mb * 16 + 4 * (lane >> 4) + i;
                if (v >= a.nvox) continue;
                const int n = nb * 16 + (lane & 15);
                float val = acc[mb][nb][i] + (a.bias ? a.bias[n] : 0.f);
                if (a.relu) val = val > 0.f ? val : 0.f;
                const long o = v * cout + n;
                if (a.mask && !(bf2f(a.mask[o]) > 0.f)) val = 0.f;
                a.out[o] = f2bf(val);
            }
}

// ---------------------------------------------------------------------------
// backward weights: dW[kk][n] += sum_v X[v + off(t)][c] * dY[v][n],
// kk = t * cin + c.  Each wave owns KBW 16-row blocks of kk and all n; a
// workgroup sweeps one voxel range and adds its partial with FP32 atomics.
constexpr int KBW = 2;

struct ConvW {
    Src s;
    int nx, ny, nz;
    long nvox;
    const bf16_t *gy;   // [nvox][NB*16]
    float *dW;          // [27 * cin][NB*16]
    int kblocks;        // 27 * cin / 16
    long vsplit;        // voxels per workgroup row of the grid (multiple of 32)
};

template <int NB>
__global__ void __launch_bounds__(256) k_conv_wgrad(ConvW a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cin = a.s.Ca + a.s.Cb, cout = NB * 16;
    const int kb0 = (blockIdx.x * 4 + wave) * KBW;
    const long vbeg = (long)blockIdx.y * a.vsplit;
    const long vend = vbeg + a.vsplit < a.nvox ? vbeg + a.vsplit : a.nvox;
    int tk[KBW], ck[KBW];
    bool kok[KBW];
#pragma unroll
    for (int kb = 0; kb < KBW; kb++) {
        kok[kb] = kb0 + kb < a.kblocks;
        const int kk = (kb0 + kb) * 16 + (lane & 15);
        tk[kb] = kk / cin;
        ck[kb] = kk - tk[kb] * cin;
    }
    v4f acc[KBW][NB];
#pragma unroll
    for (int kb = 0; kb < KBW; kb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[kb][nb] = (v4f){0.f, 0.f, 0.f, 0.f};
    for (long vs = vbeg; vs < vend; vs += 32) {
        bf16_t bv[NB][8], av[KBW][8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const long v = vs + 8 * (lane >> 4) + i;
            const bool ok = v < vend;
            int x, y, z;
            vox_xyz(ok ? v : 0, a.nx, a.ny, x, y, z);
#pragma unroll
            for (int nb = 0; nb < NB; nb++) bv[nb][i] = ok ? a.gy[v * cout + nb * 16 + (lane & 15)] : (bf16_t)0;
#pragma unroll
            for (int kb = 0; kb < KBW; kb++)
                av[kb][i] = (ok && kok[kb]) ? gather1(a.s, tk[kb], ck[kb], x, y, z, a.nx, a.ny, a.nz) : (bf16_t)0;
        }
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            v4u bw;
            bw.x = bv[nb][0] | ((uint32_t)bv[nb][1] << 16);
            bw.y = bv[nb][2] | ((uint32_t)bv[nb][3] << 16);
            bw.z = bv[nb][4] | ((uint32_t)bv[nb][5] << 16);
            bw.w = bv[nb][6] | ((uint32_t)bv[nb][7] << 16);
#pragma unroll
            for (int kb = 0; kb < KBW; kb++) {
                v4u aw;
                aw.x = av[kb][0] | ((uint32_t)av[kb][1] << 16);
                aw.y = av[kb][2] | ((uint32_t)av[kb][3] << 16);
                aw.z = av[kb][4] | ((uint32_t)av[kb][5] << 16);
                aw.w = av[kb][6] | ((uint32_t)av[kb][7] << 16);
                acc[kb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, aw),
                                                                      __builtin_bit_cast(v8bf, bw), acc[kb][nb], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int kb = 0; kb < KBW; kb++) {
        if (!kok[kb]) continue;
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int kk = (kb0 + kb) * 16 + 4 * (lane >> 4) + i;
                const int n = nb * 16 + (lane & 15);
                atomicAdd(&a.dW[(long)kk * cout + n], acc[kb][nb][i]);
            }
    }
}

// db[n] += sum_v g[v][n]
__global__ void __launch_bounds__(256) k_bias_grad(const bf16_t *__restrict__ g, long nvox, int cout, float *db) {
    __shared__ float sh[256];
    const int n = threadIdx.x % cout, lanes = 256 / cout, r = threadIdx.x / cout;
    float s = 0.f;
    for (long v = (long)blockIdx.x * lanes + r; v < nvox; v += (long)gridDim.x * lanes) s += bf2f(g[v * cout + n]);
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cout) {
        float tot = 0.f;
        for (int q = 0; q < lanes; q++) tot += sh[q * cout + threadIdx.x];
        atomicAdd(&db[threadIdx.x], tot);
    }
}

// ---------------------------------------------------------------------------
// pooling / unpooling (8 channels per thread)

// out[X][c] = mean of the 8 children of X
__global__ void k_pool(const bf16_t *__restrict__ in, bf16_t *__restrict__ out, int nx, int ny, int nz, int C) {
    const int cg = C / 8;
    const long nc = (long)(nx / 2) * (ny / 2) * (nz / 2) * cg;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nc; q += (long)gridDim.x * blockDim.x) {
        const int g = (int)(q % cg);
        const long V = q / cg;
        int X, Y, Z;
        vox_xyz(V, nx / 2, ny / 2, X, Y, Z);
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int d = 0; d < 8; d++) {
            const long v = ((long)(2 * Z + (d >> 2)) * ny + (2 * Y + ((d >> 1) & 1))) * nx + 2 * X + (d & 1);
            const v4u w = *(const v4u *)(in + v * C + g * 8);
            const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
            for (int e = 0; e < 4; e++) {
                s[2 * e] += bf2f((bf16_t)(wv[e] & 0xffff));
                s[2 * e + 1] += bf2f((bf16_t)(wv[e] >> 16));
            }
        }
        v4u o;
        uint32_t ov[4];
        for (int e = 0; e < 4; e++) ov[e] = f2bf(0.125f * s[2 * e]) | ((uint32_t)f2bf(0.125f * s[2 * e + 1]) << 16);
        o.x = ov[0]; o.y = ov[1]; o.z = ov[2]; o.w = ov[3];
        *(v4u *)(out + V * C + g * 8) = o;
    }
}

// g_fine[x][c] = mask(x,c) ? g_fine[x][c] + g_coarse[X][c] / 8 : 0
__global__ void k_pool_bwd(const bf16_t *__restrict__ gc, bf16_t *__restrict__ gf, const bf16_t *__restrict__ mask,
                           int nx, int ny, int nz, int C) {
    const long n = (long)nx * ny * nz * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long v = q / C;
        int x, y, z;
        vox_xyz(v, nx, ny, x, y, z);
        const long V = ((long)(z >> 1) * (ny >> 1) + (y >> 1)) * (nx >> 1) + (x >> 1);
        float val = bf2f(gf[q]) + 0.125f * bf2f(gc[V * C + c]);
        if (!(bf2f(mask[q]) > 0.f)) val = 0.f;
        gf[q] = f2bf(val);
    }
}

// gradient of the fused [up(A), B] input: gA[X][c] = mask ? sum of the 8
// children du[x][c] : 0 (c < Ca);  gB[x][c] = du[x][Ca + c]
__global__ void k_up_bwd(const bf16_t *__restrict__ du, int Ca, int Cb, bf16_t *__restrict__ gA,
                         const bf16_t *__restrict__ maskA, int nx, int ny, int nz) {
    const int C = Ca + Cb;
    const long n = (long)(nx / 2) * (ny / 2) * (nz / 2) * Ca;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % Ca);
        const long V = q / Ca;
        int X, Y, Z;
        vox_xyz(V, nx / 2, ny / 2, X, Y, Z);
        float s = 0.f;
        for (int d = 0; d < 8; d++) {
            const long v = ((long)(2 * Z + (d >> 2)) * ny + (2 * Y + ((d >> 1) & 1))) * nx + 2 * X + (d & 1);
            s += bf2f(du[v * C + c]);
        }
        if (!(bf2f(maskA[q]) > 0.f)) s = 0.f;
        gA[q] = f2bf(s);
    }
}

__global__ void k_take_channels(const bf16_t *__restrict__ du, int C, int c0, int Cb, bf16_t *__restrict__ gB, long nvox) {
    const long n = nvox * Cb;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % Cb);
        const long v = q / Cb;
        gB[q] = du[v * C + c0 + c];
    }
}

__global__ void k_relu_mask(bf16_t *__restrict__ g, const bf16_t *__restrict__ y, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        if (!(bf2f(y[q]) > 0.f)) g[q] = 0;
}

// ---------------------------------------------------------------------------
// 16 -> 1 head (FP32 weights rounded to BF16 like every other layer)

__device__ inline void unpack16(const bf16_t *p, float (&v)[16]) {
    const v4u a = *(const v4u *)p, b = *(const v4u *)(p + 8);
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int e = 0; e < 8; e++) {
        v[2 * e] = bf2f((bf16_t)(w[e] & 0xffff));
        v[2 * e + 1] = bf2f((bf16_t)(w[e] >> 16));
    }
}

// the head's 432 weights rounded to BF16, once per workgroup
__device__ inline void load_head_weights(const float *__restrict__ W, float *w) {
    for (int q = threadIdx.x; q < 432; q += blockDim.x) w[q] = bf2f(f2bf(W[q]));
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_head_fwd(const bf16_t *__restrict__ y0, const float *__restrict__ W,
                                                  const float *__restrict__ b, int nx, int ny, int nz,
                                                  float *__restrict__ out) {
    __shared__ float w[432];
    load_head_weights(W, w);
    const float bias = b[0];
    const long n = (long)nx * ny * nz;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, nx, ny, x, y, z);
        float s = 0.f;
        for (int t = 0; t < 27; t++) {
            int xx, yy, zz;
            if (!tap_neighbour(t, x, y, z, nx, ny, nz, xx, yy, zz)) continue;
            float a[16];
            unpack16(y0 + (((long)zz * ny + yy) * nx + xx) * 16, a);
#pragma unroll
            for (int c = 0; c < 16; c++) s += w[t * 16 + c] * a[c];
        }
        out[v] = s + bias;
    }
}

// g_y0[v][c] = (y0 > 0) ? sum_t W[t][c] * g_out[v - off(t)] : 0
__global__ void __launch_bounds__(256) k_head_bwd_data(const float *__restrict__ gout, const float *__restrict__ W,
                                                       const bf16_t *__restrict__ y0, int nx, int ny, int nz,
                                                       bf16_t *__restrict__ gy0) {
    __shared__ float w[432];
    load_head_weights(W, w);
    const long n = (long)nx * ny * nz;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, nx, ny, x, y, z);
        float s[16];
        for (int c = 0; c < 16; c++) s[c] = 0.f;
        for (int t = 0; t < 27; t++) {
            int xx, yy, zz;
            // the output voxel whose tap t reads v is v - off(t) = tap 26 - t of v
            if (!tap_neighbour(26 - t, x, y, z, nx, ny, nz, xx, yy, zz)) continue;
            const float g = gout[((long)zz * ny + yy) * nx + xx];
#pragma unroll
            for (int c = 0; c < 16; c++) s[c] += w[t * 16 + c] * g;
        }
        float m[16];
        unpack16(y0 + v * 16, m);
        for (int c = 0; c < 16; c++) gy0[v * 16 + c] = f2bf(m[c] > 0.f ? s[c] : 0.f);
    }
}

// dW[t][c] += sum_v y0[v + off(t)][c] * g_out[v] ; db += sum_v g_out[v]
// 448 threads: thread q < 432 owns (t = q / 16, c = q % 16), q == 432 the bias
__global__ void __launch_bounds__(448) k_head_wgrad(const bf16_t *__restrict__ y0, const float *__restrict__ gout,
                                                    int nx, int ny, int nz, long vsplit, float *dW, float *db) {
    const int q = threadIdx.x;
    const long n = (long)nx * ny * nz;
    const long vb = (long)blockIdx.x * vsplit, ve = vb + vsplit < n ? vb + vsplit : n;
    const int t = q / 16, c = q % 16;
    float s = 0.f;
    for (long v = vb; v < ve; v++) {
        const float g = gout[v];
        if (q < 432) {
            int x, y, z, xx, yy, zz;
            vox_xyz(v, nx, ny, x, y, z);
            if (tap_neighbour(t, x, y, z, nx, ny, nz, xx, yy, zz))
                s += bf2f(y0[(((long)zz * ny + yy) * nx + xx) * 16 + c]) * g;
        } else if (q == 432) {
            s += g;
        }
    }
    if (q < 432) atomicAdd(&dW[q], s);
    else if (q == 432) atomicAdd(db, s);
}

// loss = mean (out - target)^2 ; g_out = 2 (out - target) / n
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ out, const float *__restrict__ tgt, long n,
                                              float *__restrict__ gout, double *part) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    const float sc = 2.0f / (float)n;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        const float d = out[v] - tgt[v];
        if (gout) gout[v] = sc * d;
        acc[0] += (double)d * d;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void k_sum_blocks(const double *part, int n, double *out) {
    __shared__ double sh[16];
    double acc[1] = {0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc[0] += part[i];
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) *out = acc[0];
}

// ---------------------------------------------------------------------------
// weights: fragments for the matrix cores, Adam

// frag[((ch*NB + nb)*64 + lane)*8 + i] = Wmat[kk = ch*32 + 8*(lane/16) + i][n = nb*16 + lane%16]
// forward: Wmat[t*cin + c][n] = W[t][c][n]; flipped: Wmat[t*cout + n][c] = W[26-t][c][n]
__global__ void k_pack_frag(const float *__restrict__ W, int cin, int cout, int flipped, int nchunks,
                            bf16_t *__restrict__ frag) {
    const int kc = flipped ? cout : cin;      // K channels of the packed operand
    const int nn = flipped ? cin : cout;      // N of the packed operand
    const int NB = nn / 16;
    const long total = (long)nchunks * NB * 64 * 8;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % 8);
        const int lane = (int)((q / 8) % 64);
        const long r = q / 512;
        const int nb = (int)(r % NB);
        const int ch = (int)(r / NB);
        const int kk = ch * 32 + 8 * (lane / 16) + i;
        const int n = nb * 16 + lane % 16;
        const int t = kk / kc, k = kk % kc;
        float w = 0.f;
        if (t < 27) w = flipped ? W[((long)(26 - t) * cin + n) * cout + k] : W[((long)t * cin + k) * cout + n];
        frag[q] = f2bf(w);
    }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.f - b1) * gq;
        const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    }
}

// ---------------------------------------------------------------------------
// data movement between the solver's FP64 fields and the network

// in16[v] = (b / s_b, phi1 / s_p, phi2 / s_p, 0 ...) at cell v (x fastest)
// s_b = sqrt(ss[0] / ncells), s_p = sqrt(ss[1] / ncells) (1 when zero)
__device__ inline double inv_rms(double ss, long ncells) {
    return ss > 0.0 ? 1.0 / sqrt(ss / (double)ncells) : 1.0;
}

__global__ void k_pack_input(Dims d, int ncx, int ncy, int ncz, const double *__restrict__ b,
                             const double *__restrict__ p1, const double *__restrict__ p2, const double *__restrict__ ss,
                             bf16_t *__restrict__ in16) {
    const long n = (long)ncx * ncy * ncz;
    const double ib = inv_rms(ss[0], n), ip = inv_rms(ss[1], n);
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, ncx, ncy, x, y, z);
        const long id = gidx(d, x, y, z);
        v4u lo, hi = {0u, 0u, 0u, 0u};
        lo.x = f2bf((float)(b[id] * ib)) | ((uint32_t)f2bf((float)(p1[id] * ip)) << 16);
        lo.y = f2bf((float)(p2[id] * ip));
        lo.z = 0u;
        lo.w = 0u;
        *(v4u *)(in16 + v * 16) = lo;
        *(v4u *)(in16 + v * 16 + 8) = hi;
    }
}

// guess = phi1 + s_p * out on cells (only s_p * out when correction_only)
__global__ void k_unpack_guess(Dims d, int ncx, int ncy, int ncz, const float *__restrict__ out,
                               const double *__restrict__ p1, const double *__restrict__ ss, int correction_only,
                               double *__restrict__ guess) {
    const long n = (long)ncx * ncy * ncz;
    const double sp = 1.0 / inv_rms(ss[1], n);
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, ncx, ncy, x, y, z);
        const long id = gidx(d, x, y, z);
        guess[id] = (correction_only ? 0.0 : p1[id]) + sp * (double)out[v];
    }
}

// snapshot: in4[v] = (b/s_b, phi1/s_p, phi2/s_p, 0), tgt[v] = (phin - phi1)/s_p
__global__ void k_snap_pack(Dims d, int ncx, int ncy, int ncz, const double *__restrict__ b,
                            const double *__restrict__ p1, const double *__restrict__ p2,
                            const double *__restrict__ pn, const double *__restrict__ ss, bf16_t *__restrict__ in4,
                            float *__restrict__ tgt) {
    const long n = (long)ncx * ncy * ncz;
    const double ib = inv_rms(ss[0], n), ip = inv_rms(ss[1], n);
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, ncx, ncy, x, y, z);
        const long id = gidx(d, x, y, z);
        in4[v * 4 + 0] = f2bf((float)(b[id] * ib));
        in4[v * 4 + 1] = f2bf((float)(p1[id] * ip));
        in4[v * 4 + 2] = f2bf((float)(p2[id] * ip));
        in4[v * 4 + 3] = 0;
        tgt[v] = (float)((pn[id] - p1[id]) * ip);
    }
}

// training crop of a snapshot (sx, sy, sz) at origin (ox, oy, oz)
__global__ void k_crop(const bf16_t *__restrict__ in4, const float *__restrict__ tgt, int sx, int sy, int ox, int oy,
                       int oz, int cx, int cy, int cz, bf16_t *__restrict__ in16, float *__restrict__ tc) {
    const long n = (long)cx * cy * cz;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        vox_xyz(v, cx, cy, x, y, z);
        const long s = ((long)(z + oz) * sy + (y + oy)) * sx + (x + ox);
        v4u lo = {0u, 0u, 0u, 0u}, hi = {0u, 0u, 0u, 0u};
        lo.x = in4[s * 4] | ((uint32_t)in4[s * 4 + 1] << 16);
        lo.y = in4[s * 4 + 2];
        *(v4u *)(in16 + v * 16) = lo;
        *(v4u *)(in16 + v * 16 + 8) = hi;
        tc[v] = tgt[s];
    }
}

// dense FP32 host-side copy of a 3-channel input into in16
__global__ void k_expand3(const float *__restrict__ in3, long n, bf16_t *__restrict__ in16) {
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        v4u lo = {0u, 0u, 0u, 0u}, hi = {0u, 0u, 0u, 0u};
        lo.x = f2bf(in3[v * 3]) | ((uint32_t)f2bf(in3[v * 3 + 1]) << 16);
        lo.y = f2bf(in3[v * 3 + 2]);
        *(v4u *)(in16 + v * 16) = lo;
        *(v4u *)(in16 + v * 16 + 8) = hi;
    }
}

// ---------------------------------------------------------------------------
// host side

struct Acts {
    int nx = 0, ny = 0, nz = 0;
    long v0 = 0, v1 = 0, v2 = 0;
    bool train = false;
    bf16_t *in16 = nullptr, *x0 = nullptr, *p0 = nullptr, *x1 = nullptr, *x1b = nullptr, *p1 = nullptr,
           *x2 = nullptr, *x2b = nullptr, *y1 = nullptr, *y0 = nullptr;
    float *out = nullptr, *tgt = nullptr, *gout = nullptr;
    bf16_t *gy0 = nullptr, *du0 = nullptr, *gy1 = nullptr, *gx0 = nullptr, *du1 = nullptr, *gx2b = nullptr,
           *gx1b = nullptr, *gx2 = nullptr, *gp1 = nullptr, *gx1 = nullptr, *gp0 = nullptr;
    std::vector<void *> owned;
};

struct tb_net {
    std::string err;
    hipStream_t s = nullptr;
    long np = 0;
    long woff[NLAYER], boff[NLAYER];
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16_t *frag[NLAYER - 1] = {}, *fragT[NLAYER - 1] = {};
    int nch[NLAYER - 1], nchT[NLAYER - 1];
    int adam_t = 0;
    double *red = nullptr;      // block partials
    double *dscal = nullptr;    // device scalars
    double *hscal = nullptr;    // pinned host mirror
    Acts inf, trn;
    // snapshot store
    int cap = 0, count = 0, sx = 0, sy = 0, sz = 0;
    bf16_t *s_in = nullptr;
    float *s_tgt = nullptr;
};

#define NET_CHECK(net, call)                                                       \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            (net)->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return -1;                                                             \
        }                                                                          \
    } while (0)
#define NET_LAUNCH(net)                                                            \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            (net)->err = std::string("launch: ") + hipGetErrorString(e_);         \
            return -1;                                                             \
        }                                                                          \
    } while (0)

static const int NET_RED_BLOCKS = 1024;

static long param_layout(long *woff, long *boff) {
    long o = 0;
    for (int l = 0; l < NLAYER; l++) {
        woff[l] = o;
        o += 27L * LAYERS[l].cin * LAYERS[l].cout;
        boff[l] = o;
        o += LAYERS[l].cout;
    }
    return o;
}

static inline dim3 grid1d(long n, int bs = 256, long cap = 65536) {
    long b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return dim3((unsigned)b);
}

static int net_alloc(tb_net *net, Acts &A, void **p, size_t bytes) {
    NET_CHECK(net, hipMalloc(p, bytes));
    A.owned.push_back(*p);
    return 0;
}

static void acts_free(Acts &A) {
    for (void *p : A.owned) hipFree(p);
    A = Acts();
}

static int acts_ensure(tb_net *net, Acts &A, int nx, int ny, int nz, bool train) {
    if (A.nx == nx && A.ny == ny && A.nz == nz && (A.train || !train)) return 0;
    if (nx % 4 || ny % 4 || nz % 4 || nx < 4 || ny < 4 || nz < 4) {
        net->err = "network grid dimensions must be multiples of 4";
        return -1;
    }
    acts_free(A);
    A.nx = nx; A.ny = ny; A.nz = nz; A.train = train;
    A.v0 = (long)nx * ny * nz;
    A.v1 = A.v0 / 8;
    A.v2 = A.v1 / 8;
    const size_t b2 = sizeof(bf16_t);
    struct { bf16_t **p; long n; } bf[] = {
        {&A.in16, A.v0 * 16}, {&A.x0, A.v0 * 16}, {&A.p0, A.v1 * 16}, {&A.x1, A.v1 * 32}, {&A.x1b, A.v1 * 32},
        {&A.p1, A.v2 * 32},   {&A.x2, A.v2 * 64}, {&A.x2b, A.v2 * 64}, {&A.y1, A.v1 * 32}, {&A.y0, A.v0 * 16}};
    for (auto &e : bf)
        if (net_alloc(net, A, (void **)e.p, b2 * e.n)) return -1;
    if (net_alloc(net, A, (void **)&A.out, sizeof(float) * A.v0)) return -1;
    if (train) {
        struct { bf16_t **p; long n; } gb[] = {
            {&A.gy0, A.v0 * 16}, {&A.du0, A.v0 * 48}, {&A.gy1, A.v1 * 32}, {&A.gx0, A.v0 * 16},
            {&A.du1, A.v1 * 96}, {&A.gx2b, A.v2 * 64}, {&A.gx1b, A.v1 * 32}, {&A.gx2, A.v2 * 64},
            {&A.gp1, A.v2 * 32}, {&A.gx1, A.v1 * 32}, {&A.gp0, A.v1 * 16}};
        for (auto &e : gb)
            if (net_alloc(net, A, (void **)e.p, b2 * e.n)) return -1;
        if (net_alloc(net, A, (void **)&A.tgt, sizeof(float) * A.v0)) return -1;
        if (net_alloc(net, A, (void **)&A.gout, sizeof(float) * A.v0)) return -1;
    }
    return 0;
}

static int pack_frags(tb_net *net) {
    for (int l = 0; l < NLAYER - 1; l++) {
        const int ci = LAYERS[l].cin, co = LAYERS[l].cout;
        hipLaunchKernelGGL(k_pack_frag, grid1d((long)net->nch[l] * (co / 16) * 512), dim3(256), 0, net->s,
                           (const float *)(net->w + net->woff[l]), ci, co, 0, net->nch[l], net->frag[l]);
        hipLaunchKernelGGL(k_pack_frag, grid1d((long)net->nchT[l] * (ci / 16) * 512), dim3(256), 0, net->s,
                           (const float *)(net->w + net->woff[l]), ci, co, 1, net->nchT[l], net->fragT[l]);
    }
    NET_LAUNCH(net);
    return 0;
}

// one convolution through the matrix cores; cout selects the template
static int conv(tb_net *net, hipStream_t s, const Src &src, int nx, int ny, int nz, const bf16_t *frag, int nchunks,
                int cout, const float *bias, int relu, const bf16_t *mask, bf16_t *out) {
    ConvFwd a;
    a.s = src;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.nvox = (long)nx * ny * nz;
    a.frag = (const v4u *)frag;
    a.nchunks = nchunks;
    a.bias = bias;
    a.mask = mask;
    a.relu = relu;
    a.out = out;
    const dim3 grid((unsigned)((a.nvox + 127) / 128));
    switch (cout) {
        case 16: hipLaunchKernelGGL(k_conv_fwd<1>, grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL(k_conv_fwd<2>, grid, dim3(256), 0, s, a); break;
        case 48: hipLaunchKernelGGL(k_conv_fwd<3>, grid, dim3(256), 0, s, a); break;
        case 64: hipLaunchKernelGGL(k_conv_fwd<4>, grid, dim3(256), 0, s, a); break;
        case 96: hipLaunchKernelGGL(k_conv_fwd<6>, grid, dim3(256), 0, s, a); break;
        default: net->err = "unsupported channel count"; return -1;
    }
    NET_LAUNCH(net);
    return 0;
}

static int wgrad(tb_net *net, hipStream_t s, const Src &src, int nx, int ny, int nz, const bf16_t *gy, int cout,
                 float *dW, float *db) {
    ConvW a;
    a.s = src;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.nvox = (long)nx * ny * nz;
    a.gy = gy;
    a.dW = dW;
    const int cin = src.Ca + src.Cb;
    a.kblocks = 27 * cin / 16;
    const int gx = (a.kblocks + 4 * KBW - 1) / (4 * KBW);
    long splits = 2048 / gx;
    if (splits < 1) splits = 1;
    long maxs = (a.nvox + 255) / 256;
    if (splits > maxs) splits = maxs;
    a.vsplit = ((a.nvox + splits - 1) / splits + 31) / 32 * 32;
    splits = (a.nvox + a.vsplit - 1) / a.vsplit;
    const dim3 grid((unsigned)gx, (unsigned)splits);
    switch (cout) {
        case 16: hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL(k_conv_wgrad<2>, grid, dim3(256), 0, s, a); break;
        case 64: hipLaunchKernelGGL(k_conv_wgrad<4>, grid, dim3(256), 0, s, a); break;
        default: net->err = "unsupported channel count"; return -1;
    }
    hipLaunchKernelGGL(k_bias_grad, grid1d(a.nvox, 256, 512), dim3(256), 0, s, gy, a.nvox, cout, db);
    NET_LAUNCH(net);
    return 0;
}

static Src src1(const bf16_t *A, int C) {
    Src s;
    s.A = A; s.Ca = C; s.a_up = 0; s.B = nullptr; s.Cb = 0;
    return s;
}
static Src src2(const bf16_t *A, int Ca, const bf16_t *B, int Cb) {
    Src s;
    s.A = A; s.Ca = Ca; s.a_up = 1; s.B = B; s.Cb = Cb;
    return s;
}

static int forward(tb_net *net, hipStream_t s, Acts &A) {
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int mx = nx / 2, my = ny / 2, mz = nz / 2, qx = nx / 4, qy = ny / 4, qz = nz / 4;
    const float *W = net->w;
    auto B = [&](int l) { return W + net->boff[l]; };
    if (conv(net, s, src1(A.in16, 16), nx, ny, nz, net->frag[L_IN], net->nch[L_IN], 16, B(L_IN), 1, nullptr, A.x0)) return -1;
    hipLaunchKernelGGL(k_pool, grid1d(A.v1 * 2), dim3(256), 0, s, (const bf16_t *)A.x0, A.p0, nx, ny, nz, 16);
    if (conv(net, s, src1(A.p0, 16), mx, my, mz, net->frag[L_E1], net->nch[L_E1], 32, B(L_E1), 1, nullptr, A.x1)) return -1;
    if (conv(net, s, src1(A.x1, 32), mx, my, mz, net->frag[L_E1B], net->nch[L_E1B], 32, B(L_E1B), 1, nullptr, A.x1b)) return -1;
    hipLaunchKernelGGL(k_pool, grid1d(A.v2 * 4), dim3(256), 0, s, (const bf16_t *)A.x1b, A.p1, mx, my, mz, 32);
    if (conv(net, s, src1(A.p1, 32), qx, qy, qz, net->frag[L_E2], net->nch[L_E2], 64, B(L_E2), 1, nullptr, A.x2)) return -1;
    if (conv(net, s, src1(A.x2, 64), qx, qy, qz, net->frag[L_E2B], net->nch[L_E2B], 64, B(L_E2B), 1, nullptr, A.x2b)) return -1;
    if (conv(net, s, src2(A.x2b, 64, A.x1b, 32), mx, my, mz, net->frag[L_D1], net->nch[L_D1], 32, B(L_D1), 1, nullptr,
             A.y1)) return -1;
    if (conv(net, s, src2(A.y1, 32, A.x0, 16), nx, ny, nz, net->frag[L_D0], net->nch[L_D0], 16, B(L_D0), 1, nullptr,
             A.y0)) return -1;
    hipLaunchKernelGGL(k_head_fwd, grid1d(A.v0), dim3(256), 0, s, (const bf16_t *)A.y0, W + net->woff[L_OUT],
                       B(L_OUT), nx, ny, nz, A.out);
    NET_LAUNCH(net);
    return 0;
}

// loss of A.out against A.tgt; g_out when grad; result in *loss (host, synced)
static int loss_eval(tb_net *net, hipStream_t s, Acts &A, bool grad, double *loss) {
    hipLaunchKernelGGL(k_loss, dim3(NET_RED_BLOCKS), dim3(256), 0, s, (const float *)A.out, (const float *)A.tgt, A.v0,
                       grad ? A.gout : nullptr, net->red);
    hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(256), 0, s, (const double *)net->red, NET_RED_BLOCKS, net->dscal);
    NET_LAUNCH(net);
    if (loss) {
        NET_CHECK(net, hipMemcpyAsync(net->hscal, net->dscal, sizeof(double), hipMemcpyDeviceToHost, s));
        NET_CHECK(net, hipStreamSynchronize(s));
        *loss = net->hscal[0] / (double)A.v0;
    }
    return 0;
}

// gradients of the mean-square loss into net->g (accumulated)
static int backward(tb_net *net, hipStream_t s, Acts &A) {
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int mx = nx / 2, my = ny / 2, mz = nz / 2, qx = nx / 4, qy = ny / 4, qz = nz / 4;
    float *G = net->g;
    auto GW = [&](int l) { return G + net->woff[l]; };
    auto GB = [&](int l) { return G + net->boff[l]; };
    // head
    hipLaunchKernelGGL(k_head_bwd_data, grid1d(A.v0), dim3(256), 0, s, (const float *)A.gout,
                       (const float *)(net->w + net->woff[L_OUT]), (const bf16_t *)A.y0, nx, ny, nz, A.gy0);
    {
        long vs = (A.v0 + 255) / 256;
        if (vs < 64) vs = 64;
        const unsigned nb = (unsigned)((A.v0 + vs - 1) / vs);
        hipLaunchKernelGGL(k_head_wgrad, dim3(nb), dim3(448), 0, s, (const bf16_t *)A.y0, (const float *)A.gout, nx, ny,
                           nz, vs, GW(L_OUT), GB(L_OUT));
    }
    NET_LAUNCH(net);
    // y0 = relu(conv([up(y1), x0]))  (gy0 already masked)
    if (wgrad(net, s, src2(A.y1, 32, A.x0, 16), nx, ny, nz, A.gy0, 16, GW(L_D0), GB(L_D0))) return -1;
    if (conv(net, s, src1(A.gy0, 16), nx, ny, nz, net->fragT[L_D0], net->nchT[L_D0], 48, nullptr, 0, nullptr, A.du0)) return -1;
    hipLaunchKernelGGL(k_up_bwd, grid1d(A.v1 * 32), dim3(256), 0, s, (const bf16_t *)A.du0, 32, 16, A.gy1,
                       (const bf16_t *)A.y1, nx, ny, nz);
    hipLaunchKernelGGL(k_take_channels, grid1d(A.v0 * 16), dim3(256), 0, s, (const bf16_t *)A.du0, 48, 32, 16, A.gx0, A.v0);
    NET_LAUNCH(net);
    // y1 = relu(conv([up(x2b), x1b]))
    if (wgrad(net, s, src2(A.x2b, 64, A.x1b, 32), mx, my, mz, A.gy1, 32, GW(L_D1), GB(L_D1))) return -1;
    if (conv(net, s, src1(A.gy1, 32), mx, my, mz, net->fragT[L_D1], net->nchT[L_D1], 96, nullptr, 0, nullptr, A.du1)) return -1;
    hipLaunchKernelGGL(k_up_bwd, grid1d(A.v2 * 64), dim3(256), 0, s, (const bf16_t *)A.du1, 64, 32, A.gx2b,
                       (const bf16_t *)A.x2b, mx, my, mz);
    hipLaunchKernelGGL(k_take_channels, grid1d(A.v1 * 32), dim3(256), 0, s, (const bf16_t *)A.du1, 96, 64, 32, A.gx1b, A.v1);
    NET_LAUNCH(net);
    // x2b = relu(conv(x2))
    if (wgrad(net, s, src1(A.x2, 64), qx, qy, qz, A.gx2b, 64, GW(L_E2B), GB(L_E2B))) return -1;
    if (conv(net, s, src1(A.gx2b, 64), qx, qy, qz, net->fragT[L_E2B], net->nchT[L_E2B], 64, nullptr, 0, A.x2, A.gx2)) return -1;
    // x2 = relu(conv(p1))
    if (wgrad(net, s, src1(A.p1, 32), qx, qy, qz, A.gx2, 64, GW(L_E2), GB(L_E2))) return -1;
    if (conv(net, s, src1(A.gx2, 64), qx, qy, qz, net->fragT[L_E2], net->nchT[L_E2], 32, nullptr, 0, nullptr, A.gp1)) return -1;
    // p1 = pool(x1b): gx1b += up(gp1)/8, then the x1b ReLU mask
    hipLaunchKernelGGL(k_pool_bwd, grid1d(A.v1 * 32), dim3(256), 0, s, (const bf16_t *)A.gp1, A.gx1b,
                       (const bf16_t *)A.x1b, mx, my, mz, 32);
    NET_LAUNCH(net);
    // x1b = relu(conv(x1))
    if (wgrad(net, s, src1(A.x1, 32), mx, my, mz, A.gx1b, 32, GW(L_E1B), GB(L_E1B))) return -1;
    if (conv(net, s, src1(A.gx1b, 32), mx, my, mz, net->fragT[L_E1B], net->nchT[L_E1B], 32, nullptr, 0, A.x1, A.gx1)) return -1;
    // x1 = relu(conv(p0))
    if (wgrad(net, s, src1(A.p0, 16), mx, my, mz, A.gx1, 32, GW(L_E1), GB(L_E1))) return -1;
    if (conv(net, s, src1(A.gx1, 32), mx, my, mz, net->fragT[L_E1], net->nchT[L_E1], 16, nullptr, 0, nullptr, A.gp0)) return -1;
    // p0 = pool(x0)
    hipLaunchKernelGGL(k_pool_bwd, grid1d(A.v0 * 16), dim3(256), 0, s, (const bf16_t *)A.gp0, A.gx0,
                       (const bf16_t *)A.x0, nx, ny, nz, 16);
    NET_LAUNCH(net);
    // x0 = relu(conv(in))
    if (wgrad(net, s, src1(A.in16, 16), nx, ny, nz, A.gx0, 16, GW(L_IN), GB(L_IN))) return -1;
    return 0;
}

static uint64_t splitmix(uint64_t &st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif01(uint64_t &st) { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0); }

// He-uniform weights (input layer: only the 3 used channels), zero biases,
// head scaled by 0.1 so the untrained network returns phi_{n-1} + O(0.1)
static void init_weights(std::vector<float> &w, const long *woff, const long *boff, uint64_t seed) {
    uint64_t st = seed;
    for (int l = 0; l < NLAYER; l++) {
        const int ci = LAYERS[l].cin, co = LAYERS[l].cout;
        const int fan = 27 * (l == L_IN ? 3 : ci);
        float a = (float)sqrt(6.0 / fan);
        if (l == L_OUT) a *= 0.1f;
        for (int t = 0; t < 27; t++)
            for (int c = 0; c < ci; c++)
                for (int n = 0; n < co; n++) {
                    float r = (float)(2.0 * unif01(st) - 1.0) * a;
                    if (l == L_IN && c >= 3) r = 0.f;
                    w[woff[l] + ((long)t * ci + c) * co + n] = r;
                }
        for (int n = 0; n < co; n++) w[boff[l] + n] = 0.f;
    }
}

extern "C" int64_t tb_net_param_count(void) {
    long wo[NLAYER], bo[NLAYER];
    return param_layout(wo, bo);
}

extern "C" tb_net *tb_net_create(uint64_t seed) {
    tb_net *net = new tb_net();
    net->np = param_layout(net->woff, net->boff);
    bool ok = hipStreamCreate(&net->s) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->w, sizeof(float) * net->np) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->g, sizeof(float) * net->np) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->m, sizeof(float) * net->np) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->v, sizeof(float) * net->np) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->red, sizeof(double) * NET_RED_BLOCKS) == hipSuccess;
    ok = ok && hipMalloc((void **)&net->dscal, sizeof(double) * 8) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&net->hscal, sizeof(double) * 8, hipHostMallocCoherent) == hipSuccess;
    for (int l = 0; ok && l < NLAYER - 1; l++) {
        const int ci = LAYERS[l].cin, co = LAYERS[l].cout;
        net->nch[l] = (27 * ci + 31) / 32;
        net->nchT[l] = (27 * co + 31) / 32;
        ok = hipMalloc((void **)&net->frag[l], sizeof(bf16_t) * 512L * net->nch[l] * (co / 16)) == hipSuccess &&
             hipMalloc((void **)&net->fragT[l], sizeof(bf16_t) * 512L * net->nchT[l] * (ci / 16)) == hipSuccess;
    }
    if (!ok) {
        tb_net_destroy(net);
        return nullptr;
    }
    std::vector<float> w(net->np);
    init_weights(w, net->woff, net->boff, seed);
    ok = hipMemcpy(net->w, w.data(), sizeof(float) * net->np, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(net->m, 0, sizeof(float) * net->np) == hipSuccess &&
         hipMemset(net->v, 0, sizeof(float) * net->np) == hipSuccess;
    if (!ok || pack_frags(net) || hipStreamSynchronize(net->s) != hipSuccess) {
        tb_net_destroy(net);
        return nullptr;
    }
    return net;
}

extern "C" void tb_net_destroy(tb_net *net) {
    if (!net) return;
    if (net->s) hipStreamSynchronize(net->s);
    acts_free(net->inf);
    acts_free(net->trn);
    hipFree(net->w); hipFree(net->g); hipFree(net->m); hipFree(net->v);
    hipFree(net->red); hipFree(net->dscal);
    if (net->hscal) hipHostFree(net->hscal);
    for (int l = 0; l < NLAYER - 1; l++) { hipFree(net->frag[l]); hipFree(net->fragT[l]); }
    hipFree(net->s_in); hipFree(net->s_tgt);
    if (net->s) hipStreamDestroy(net->s);
    delete net;
}

extern "C" const char *tb_net_last_error(tb_net *net) { return net ? net->err.c_str() : "null network"; }

extern "C" int tb_net_get_weights(tb_net *net, float *host, int64_t n) {
    if (!net || !host || n != net->np) { if (net) net->err = "tb_net_get_weights: bad size"; return -1; }
    NET_CHECK(net, hipMemcpyAsync(host, net->w, sizeof(float) * n, hipMemcpyDeviceToHost, net->s));
    NET_CHECK(net, hipStreamSynchronize(net->s));
    return 0;
}

extern "C" int tb_net_set_weights(tb_net *net, const float *host, int64_t n) {
    if (!net || !host || n != net->np) { if (net) net->err = "tb_net_set_weights: bad size"; return -1; }
    NET_CHECK(net, hipMemcpyAsync(net->w, host, sizeof(float) * n, hipMemcpyHostToDevice, net->s));
    NET_CHECK(net, hipMemsetAsync(net->m, 0, sizeof(float) * n, net->s));
    NET_CHECK(net, hipMemsetAsync(net->v, 0, sizeof(float) * n, net->s));
    net->adam_t = 0;
    if (pack_frags(net)) return -1;
    NET_CHECK(net, hipStreamSynchronize(net->s));
    return 0;
}

// host in3 [nz][ny][nx][3] FP32 -> out [nz][ny][nx] FP32
extern "C" int tb_net_forward(tb_net *net, const float *in3, int nx, int ny, int nz, float *out) {
    if (!net || !in3 || !out) return -1;
    if (acts_ensure(net, net->trn, nx, ny, nz, true)) return -1;
    Acts &A = net->trn;
    float *d3 = nullptr;
    NET_CHECK(net, hipMalloc((void **)&d3, sizeof(float) * 3 * A.v0));
    int rc = 0;
    if (hipMemcpyAsync(d3, in3, sizeof(float) * 3 * A.v0, hipMemcpyHostToDevice, net->s) != hipSuccess) rc = -1;
    if (!rc) hipLaunchKernelGGL(k_expand3, grid1d(A.v0), dim3(256), 0, net->s, (const float *)d3, A.v0, A.in16);
    if (!rc) rc = forward(net, net->s, A);
    if (!rc && hipMemcpyAsync(out, A.out, sizeof(float) * A.v0, hipMemcpyDeviceToHost, net->s) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(net->s) != hipSuccess) rc = -1;
    hipFree(d3);
    if (rc && net->err.empty()) net->err = "tb_net_forward failed";
    return rc;
}

// loss and full parameter gradient for host input / target (tests)
extern "C" int tb_net_gradient(tb_net *net, const float *in3, const float *target, int nx, int ny, int nz,
                               double *loss, float *grad) {
    if (!net || !in3 || !target || !grad) return -1;
    if (acts_ensure(net, net->trn, nx, ny, nz, true)) return -1;
    Acts &A = net->trn;
    float *d3 = nullptr;
    NET_CHECK(net, hipMalloc((void **)&d3, sizeof(float) * 3 * A.v0));
    int rc = 0;
    if (hipMemcpyAsync(d3, in3, sizeof(float) * 3 * A.v0, hipMemcpyHostToDevice, net->s) != hipSuccess ||
        hipMemcpyAsync(A.tgt, target, sizeof(float) * A.v0, hipMemcpyHostToDevice, net->s) != hipSuccess ||
        hipMemsetAsync(net->g, 0, sizeof(float) * net->np, net->s) != hipSuccess)
        rc = -1;
    if (!rc) hipLaunchKernelGGL(k_expand3, grid1d(A.v0), dim3(256), 0, net->s, (const float *)d3, A.v0, A.in16);
    if (!rc) rc = forward(net, net->s, A);
    if (!rc) rc = loss_eval(net, net->s, A, true, loss);
    if (!rc) rc = backward(net, net->s, A);
    if (!rc && hipMemcpyAsync(grad, net->g, sizeof(float) * net->np, hipMemcpyDeviceToHost, net->s) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(net->s) != hipSuccess) rc = -1;
    hipFree(d3);
    if (rc && net->err.empty()) net->err = "tb_net_gradient failed";
    return rc;
}

// snapshot store ------------------------------------------------------------

extern "C" int tb_net_reserve(tb_net *net, int capacity, int nx, int ny, int nz) {
    if (!net || capacity < 1 || nx < 4 || ny < 4 || nz < 4) { if (net) net->err = "tb_net_reserve: bad argument"; return -1; }
    hipFree(net->s_in); hipFree(net->s_tgt);
    net->s_in = nullptr; net->s_tgt = nullptr;
    const long V = (long)nx * ny * nz;
    NET_CHECK(net, hipMalloc((void **)&net->s_in, sizeof(bf16_t) * 4 * V * capacity));
    NET_CHECK(net, hipMalloc((void **)&net->s_tgt, sizeof(float) * V * capacity));
    net->cap = capacity; net->count = 0;
    net->sx = nx; net->sy = ny; net->sz = nz;
    return 0;
}

extern "C" int tb_net_snapshot_count(tb_net *net) { return net ? net->count : -1; }

// (b_n, phi_{n-1}, phi_{n-2}) -> phi_n of the context's last Poisson solve
extern "C" int tb_net_add_snapshot(tb_net *net, tb_ctx *c) {
    if (!net || !c) return -1;
    if (c->ncx != net->sx || c->ncy != net->sy || c->ncz != net->sz) { net->err = "snapshot grid differs from the store"; return -1; }
    if (c->n_hist < 3) { net->err = "need three solved pressures"; return -1; }
    if (net->count >= net->cap) { net->err = "snapshot store full"; return -1; }
    const long V = (long)net->sx * net->sy * net->sz;
    if (tb_dot(c, c->f[TB_F_PRHS], c->f[TB_F_PRHS], S_TMP1)) { net->err = c->err; return -1; }
    if (tb_dot(c, c->phi_hist[1], c->phi_hist[1], S_TMP2)) { net->err = c->err; return -1; }
    hipLaunchKernelGGL(k_snap_pack, grid1d(V), dim3(256), 0, c->s, c->d, c->ncx, c->ncy, c->ncz,
                       (const double *)c->f[TB_F_PRHS], (const double *)c->phi_hist[1], (const double *)c->phi_hist[2],
                       (const double *)c->phi_hist[0], (const double *)(c->scal + S_TMP1),
                       net->s_in + 4 * V * net->count, net->s_tgt + V * net->count);
    NET_LAUNCH(net);
    NET_CHECK(net, hipStreamSynchronize(c->s));
    net->count++;
    return 0;
}

// Adam on random crops (cx, cy, cz) of random snapshots; batch crops per step
extern "C" int tb_net_train(tb_net *net, int steps, int batch, int cx, int cy, int cz, double lr, uint64_t seed,
                            double *loss_first, double *loss_last) {
    if (!net || net->count < 1 || steps < 0 || batch < 1) { if (net) net->err = "tb_net_train: no snapshots"; return -1; }
    if (cx > net->sx || cy > net->sy || cz > net->sz) { net->err = "crop larger than the snapshots"; return -1; }
    if (acts_ensure(net, net->trn, cx, cy, cz, true)) return -1;
    Acts &A = net->trn;
    const long V = (long)net->sx * net->sy * net->sz;
    uint64_t st = seed;
    double lsum = 0.0;
    for (int it = 0; it < steps; it++) {
        NET_CHECK(net, hipMemsetAsync(net->g, 0, sizeof(float) * net->np, net->s));
        lsum = 0.0;
        for (int bi = 0; bi < batch; bi++) {
            const int sidx = (int)(splitmix(st) % (uint64_t)net->count);
            const int ox = (int)(splitmix(st) % (uint64_t)(net->sx - cx + 1));
            const int oy = (int)(splitmix(st) % (uint64_t)(net->sy - cy + 1));
            const int oz = (int)(splitmix(st) % (uint64_t)(net->sz - cz + 1));
            hipLaunchKernelGGL(k_crop, grid1d(A.v0), dim3(256), 0, net->s, (const bf16_t *)(net->s_in + 4 * V * sidx),
                               (const float *)(net->s_tgt + V * sidx), net->sx, net->sy, ox, oy, oz, cx, cy, cz, A.in16,
                               A.tgt);
            NET_LAUNCH(net);
            if (forward(net, net->s, A)) return -1;
            double l = 0.0;
            const bool report = (it == 0 || it == steps - 1);
            if (loss_eval(net, net->s, A, true, report ? &l : nullptr)) return -1;
            lsum += l;
            if (backward(net, net->s, A)) return -1;
        }
        if (it == 0 && loss_first) *loss_first = lsum / batch;
        if (it == steps - 1 && loss_last) *loss_last = lsum / batch;
        net->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)net->adam_t), c2 = 1.f - powf(b2, (float)net->adam_t);
        hipLaunchKernelGGL(k_adam, grid1d(net->np), dim3(256), 0, net->s, net->w, (const float *)net->g, net->m, net->v,
                           net->np, (float)lr, b1, b2, 1e-8f, c1, c2);
        NET_LAUNCH(net);
        if (pack_frags(net)) return -1;
    }
    NET_CHECK(net, hipStreamSynchronize(net->s));
    return 0;
}

// mean loss over whole stored snapshots [first, first + count)
extern "C" int tb_net_eval_snapshots(tb_net *net, int first, int count, double *mean_loss) {
    if (!net || first < 0 || count < 1 || first + count > net->count || !mean_loss) { if (net) net->err = "bad range"; return -1; }
    if (acts_ensure(net, net->trn, net->sx, net->sy, net->sz, true)) return -1;
    Acts &A = net->trn;
    const long V = (long)net->sx * net->sy * net->sz;
    double tot = 0.0;
    for (int k = first; k < first + count; k++) {
        hipLaunchKernelGGL(k_crop, grid1d(A.v0), dim3(256), 0, net->s, (const bf16_t *)(net->s_in + 4 * V * k),
                           (const float *)(net->s_tgt + V * k), net->sx, net->sy, 0, 0, 0, net->sx, net->sy, net->sz,
                           A.in16, A.tgt);
        NET_LAUNCH(net);
        if (forward(net, net->s, A)) return -1;
        double l;
        if (loss_eval(net, net->s, A, false, &l)) return -1;
        tot += l;
    }
    *mean_loss = tot / count;
    return 0;
}

// solver side ---------------------------------------------------------------

// every later Poisson solve of ctx appends its snapshot to net's store until
// the store is full; NULL stops
extern "C" int tb_net_record(tb_ctx *c, tb_net *net) {
    if (!c) return -1;
    if (net && (c->ncx != net->sx || c->ncy != net->sy || c->ncz != net->sz)) {
        c->err = "snapshot store not reserved for this grid";
        return -1;
    }
    c->net_record = net;
    return 0;
}

int tb_net_record_one(tb_ctx *c) {
    tb_net *net = c->net_record;
    if (net->count >= net->cap) return 0;
    if (tb_net_add_snapshot(net, c)) { c->err = net->err; return -1; }
    return 0;
}

extern "C" int tb_net_attach(tb_ctx *c, tb_net *net) {
    if (!c) return -1;
    if (net && (c->ncx % 4 || c->ncy % 4 || c->ncz % 4)) { c->err = "network needs cell counts divisible by 4"; return -1; }
    for (int e = 0; e < 2; e++)
        if (!c->net_ev[e]) TB_CHECK(c, hipEventCreate(&c->net_ev[e]));
    c->net = net;
    c->net_open = true;
    c->net_trials = 0;
    c->net_saved = c->net_cost_ms = c->net_iter_ms = 0.0;
    return 0;
}

// guess from PRHS, phi_{n-1} = hist[0], phi_{n-2} = hist[1], on the
// context's stream; cells only (the caller owns the other entries)
int tb_net_predict(tb_ctx *c, double *guess, int correction_only) {
    tb_net *net = c->net;
    if (!net || c->n_hist < 2) {
        c->err = "network guess: needs an attached network and two previous solutions";
        return -1;
    }
    if (acts_ensure(net, net->inf, c->ncx, c->ncy, c->ncz, false)) { c->err = net->err; return -1; }
    Acts &A = net->inf;
    if (tb_dot(c, c->f[TB_F_PRHS], c->f[TB_F_PRHS], S_TMP1)) return -1;
    if (tb_dot(c, c->phi_hist[0], c->phi_hist[0], S_TMP2)) return -1;
    if (c->prof_on) tb_prof_begin(c, TB_K_NET);
    hipLaunchKernelGGL(k_pack_input, grid1d(A.v0), dim3(256), 0, c->s, c->d, c->ncx, c->ncy, c->ncz,
                       (const double *)c->f[TB_F_PRHS], (const double *)c->phi_hist[0],
                       (const double *)c->phi_hist[1], (const double *)(c->scal + S_TMP1), A.in16);
    if (forward(net, c->s, A)) { c->err = net->err; return -1; }
    hipLaunchKernelGGL(k_unpack_guess, grid1d(A.v0), dim3(256), 0, c->s, c->d, c->ncx, c->ncy, c->ncz,
                       (const float *)A.out, (const double *)c->phi_hist[0], (const double *)(c->scal + S_TMP1),
                       correction_only, guess);
    if (c->prof_on) tb_prof_end(c, TB_K_NET);
    TB_LAUNCH_CHECK(c);
    return 0;
}

extern "C" int tb_net_predict_field(tb_ctx *c, int f_out) {
    if (!c || f_out < 0 || f_out >= TB_F_COUNT) return -1;
    if (!c->net) { c->err = "no network attached"; return -1; }
    if (c->n_hist < 2) { c->err = "need two solved pressures"; return -1; }
    tb_field_written(c, f_out);
    TB_CHECK(c, hipMemsetAsync(c->f[f_out], 0, sizeof(double) * c->d.ntot, c->s));
    if (tb_net_predict(c, c->f[f_out], 0)) return -1;
    TB_CHECK(c, hipStreamSynchronize(c->s));
    return 0;
}
