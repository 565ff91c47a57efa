// unet.hip -- learned Poisson warm start / preconditioner stage for gfx950.
//
// A 3D convolutional U-Net on the interior cells (ni x nj x nk, each divisible
// by 2^(levels-1)): level l has base*2^l channels; the encoder applies one
// 3x3x3 convolution + ReLU per level with 2x2x2 average pooling between
// levels, the decoder upsamples (nearest), concatenates the encoder output of
// the same level and applies one 3x3x3 convolution + ReLU; a 1x1x1
// convolution gives the single output channel.  Zero padding at the edges.
//
// Precision: BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_16x16x16_bf16), FP32 master weights, gradients and Adam
// state.  Activations are channels-last ([k][j][i][c], i fastest among the
// points) so that one lane's four consecutive channels are one 8-byte load.
//
// Convolution as implicit GEMM: M = points (a 64-point run along i per
// workgroup, 16 per wave), N = output channels (16 per MFMA tile), K = 27
// taps x input channels (16 per step).  The data gradient is the same kernel
// with the taps flipped and the weights transposed; the weight gradient is a
// second MFMA kernel with K = points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"

namespace vfs {

typedef uint16_t bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16 f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return (bf16)(u >> 16);
}

__host__ __device__ inline float bf2f(bf16 b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Dims {
    int ni, nj, nk;
    __host__ __device__ long long np() const { return (long long)ni * nj * nk; }
};

constexpr int CV_TPB = 256;  // 4 waves, 64 points along i

// Y[p][co] = act(sum_{tap, ci} X[p + tap][ci] W[tap][ci][co] + bias[co]).
// Wf holds B fragments: Wf[((tap * CB + cb) * NT + t) * 64 + lane] = the four
// weights (ci = 16 cb + 4 (lane/16) + 0..3, co = 16 t + lane%16).
template <int NT, bool RELU, bool OUT_F32>
__global__ void __launch_bounds__(CV_TPB) k_conv3(Dims d, int cin, const bf16 *__restrict__ X, const bf16x4 *__restrict__ Wf,
                                                  const float *__restrict__ bias, bool flip, void *__restrict__ Yv)
{
    const int cout = 16 * NT;
    const int CB = cin / 16;
    const int nbi = (d.ni + 63) / 64;
    const int bi = blockIdx.x % nbi;
    const int j = (blockIdx.x / nbi) % d.nj;
    const int k = blockIdx.x / (nbi * d.nj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const int i = bi * 64 + wave * 16 + m;  // this lane's A row (point)
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; tap++) {
        const int st = flip ? 26 - tap : tap;  // data gradient: mirrored stencil
        const int dk = st / 9 - 1, dj = (st / 3) % 3 - 1, di = st % 3 - 1;
        const int ii = i + di, jj = j + dj, kk = k + dk;
        const bool in = ii >= 0 && ii < d.ni && jj >= 0 && jj < d.nj && kk >= 0 && kk < d.nk;
        const long long p = in ? ((long long)kk * d.nj + jj) * d.ni + ii : 0;
        const bf16x4 *xa = reinterpret_cast<const bf16x4 *>(X + p * cin) + kq;
        const bf16x4 *wt = Wf + (long long)tap * CB * NT * 64 + lane;
        for (int cb = 0; cb < CB; cb++) {
            const bf16x4 a = in ? xa[cb * 4] : bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wt[(cb * NT + t) * 64], acc[t], 0, 0, 0);
        }
    }
    // lane holds rows 4*kq + r (points) of column m (channel) of each tile
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = 16 * t + m;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ip = bi * 64 + wave * 16 + 4 * kq + r;
            if (ip >= d.ni) continue;
            float v = acc[t][r] + bb;
            if (RELU) v = v > 0.f ? v : 0.f;
            const long long q = (((long long)k * d.nj + j) * d.ni + ip) * cout + co;
            if (OUT_F32) reinterpret_cast<float *>(Yv)[q] = v;
            else reinterpret_cast<bf16 *>(Yv)[q] = f2bf(v);
        }
    }
}

// dW[tap][ci][co] += sum_p X[p + tap][ci] G[p][co] over 16 rows of 64 points
// per wave; one wave per (row group, tap, ci tile, co tile)
constexpr int WG_ROWS = 16;
__global__ void __launch_bounds__(64) k_conv3_wgrad(Dims d, int cin, int cout, const bf16 *__restrict__ X,
                                                    const bf16 *__restrict__ G, float *__restrict__ dW)
{
    const int CB = cin / 16, NT = cout / 16;
    int w = blockIdx.y;
    const int t = w % NT;
    w /= NT;
    const int cb = w % CB, tap = w / CB;
    const int dk = tap / 9 - 1, dj = (tap / 3) % 3 - 1, di = tap % 3 - 1;
    const int lane = threadIdx.x, m = lane & 15, kq = lane >> 4;
    const int nbi = (d.ni + 63) / 64;
    const long long nrows = (long long)nbi * d.nj * d.nk;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long row = (long long)blockIdx.x * WG_ROWS; row < nrows && row < ((long long)blockIdx.x + 1) * WG_ROWS; row++) {
        const int bi = row % nbi, j = (row / nbi) % d.nj, k = row / ((long long)nbi * d.nj);
        const int jj = j + dj, kk = k + dk;
        const bool rowin = jj >= 0 && jj < d.nj && kk >= 0 && kk < d.nk;
        for (int s = 0; s < 4; s++) {  // 16 points per MFMA step
            bf16x4 a, b;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = bi * 64 + s * 16 + 4 * kq + r;  // K index = point
                const int ii = i + di;
                const bool pin = i < d.ni;
                const bool xin = rowin && pin && ii >= 0 && ii < d.ni;
                const long long px = ((long long)kk * d.nj + jj) * d.ni + ii;
                const long long pg = ((long long)k * d.nj + j) * d.ni + i;
                a[r] = xin ? (short)X[px * cin + cb * 16 + m] : (short)0;
                b[r] = pin ? (short)G[pg * cout + t * 16 + m] : (short)0;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    // acc[r] = dW[ci = 4 kq + r][co = m] for this tile
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int ci = cb * 16 + 4 * kq + r, co = t * 16 + m;
        atomicAdd(&dW[((long long)tap * cin + ci) * cout + co], acc[r]);
    }
}

// B-fragment packing of FP32 master weights W[27][cin][cout] (forward) or of
// the transposed weights W'[tap][co][ci] = W[tap][ci][co] (data gradient, the
// kernel then walks the taps mirrored)
__global__ void k_pack(int cin, int cout, int transpose, const float *__restrict__ W, bf16x4 *__restrict__ Wf)
{
    const int kin = transpose ? cout : cin, kout = transpose ? cin : cout;
    const int CB = kin / 16, NT = kout / 16;
    const long long total = 27LL * CB * NT * 64;
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int lane = q % 64;
    long long r = q / 64;
    const int t = r % NT;
    r /= NT;
    const int cb = r % CB, tap = r / CB;
    bf16x4 v;
    for (int e = 0; e < 4; e++) {
        const int a = cb * 16 + 4 * (lane >> 4) + e, o = t * 16 + (lane & 15);
        const float w = transpose ? W[((long long)tap * cin + o) * cout + a] : W[((long long)tap * cin + a) * cout + o];
        v[e] = (short)f2bf(w);
    }
    Wf[q] = v;
}

// ---- elementwise layers ------------------------------------------------------
// input channel 0 = x * scale (double -> bf16), channels 1..15 = 0
__global__ void k_in(Dims d, Grid g, const double *__restrict__ x, const double *__restrict__ scale, bf16 *__restrict__ X)
{
    long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= d.np()) return;
    const int i = p % d.ni, j = (p / d.ni) % d.nj, k = p / ((long long)d.ni * d.nj);
    const double v = x[((long long)(k + 1) * g.my + j + 1) * g.mx + i + 1] * scale[0];
    bf16 *o = X + p * 16;
    o[0] = f2bf((float)v);
    for (int c = 1; c < 16; c++) o[c] = 0;
}

// 2x2x2 average pooling, channels-last
__global__ void k_pool(Dims dc, int C, const bf16 *__restrict__ Xf, bf16 *__restrict__ Xc)
{
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= dc.np() * C) return;
    const int c = q % C;
    const long long P = q / C;
    const int I = P % dc.ni, J = (P / dc.ni) % dc.nj, K = P / ((long long)dc.ni * dc.nj);
    const int nif = 2 * dc.ni, njf = 2 * dc.nj;
    float s = 0.f;
    for (int e = 0; e < 8; e++) {
        const long long pf = ((long long)(2 * K + (e >> 2)) * njf + 2 * J + ((e >> 1) & 1)) * nif + 2 * I + (e & 1);
        s += bf2f(Xf[pf * C + c]);
    }
    Xc[q] = f2bf(s * 0.125f);
}

// gf[fine] += gc[parent] / 8
__global__ void k_pool_back(Dims df, int C, const float *__restrict__ gc, float *__restrict__ gf)
{
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= df.np() * C) return;
    const int c = q % C;
    const long long p = q / C;
    const int i = p % df.ni, j = (p / df.ni) % df.nj, k = p / ((long long)df.ni * df.nj);
    const long long P = ((long long)(k / 2) * (df.nj / 2) + j / 2) * (df.ni / 2) + i / 2;
    gf[q] += gc[P * C + c] * 0.125f;
}

// concat buffer of the decoder: [upsampled coarse (Cu) | skip (Cs)]
__global__ void k_up_cat(Dims df, int Cu, int Cs, const bf16 *__restrict__ Yc, const bf16 *__restrict__ S,
                         bf16 *__restrict__ Z)
{
    const int C = Cu + Cs;
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= df.np() * C) return;
    const int c = q % C;
    const long long p = q / C;
    const int i = p % df.ni, j = (p / df.ni) % df.nj, k = p / ((long long)df.ni * df.nj);
    if (c < Cu) {
        const long long P = ((long long)(k / 2) * (df.nj / 2) + j / 2) * (df.ni / 2) + i / 2;
        Z[q] = Yc[P * Cu + c];
    } else {
        Z[q] = S[p * Cs + c - Cu];
    }
}

// gradient of the concat: coarse part summed over the 8 children, skip part
// added to the encoder gradient
__global__ void k_up_cat_back_coarse(Dims dc, int Cu, int Cs, const float *__restrict__ gz, float *__restrict__ gc)
{
    const int C = Cu + Cs;
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= dc.np() * Cu) return;
    const int c = q % Cu;
    const long long P = q / Cu;
    const int I = P % dc.ni, J = (P / dc.ni) % dc.nj, K = P / ((long long)dc.ni * dc.nj);
    const int nif = 2 * dc.ni, njf = 2 * dc.nj;
    float s = 0.f;
    for (int e = 0; e < 8; e++) {
        const long long pf = ((long long)(2 * K + (e >> 2)) * njf + 2 * J + ((e >> 1) & 1)) * nif + 2 * I + (e & 1);
        s += gz[pf * C + c];
    }
    gc[q] = s;
}

__global__ void k_cat_back_skip(Dims df, int Cu, int Cs, const float *__restrict__ gz, float *__restrict__ gs)
{
    const int C = Cu + Cs;
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= df.np() * Cs) return;
    const int c = q % Cs;
    const long long p = q / Cs;
    gs[q] = gz[p * C + Cu + c];
}

// masked output gradient (ReLU) as BF16 for the backward convolutions
__global__ void k_relu_mask(long long n, const bf16 *__restrict__ y, const float *__restrict__ g, bf16 *__restrict__ gm)
{
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    gm[q] = bf2f(y[q]) > 0.f ? f2bf(g[q]) : (bf16)0;
}

// bias gradient: db[c] += sum_p gm[p][c]
__global__ void k_bias_grad(long long np, int C, const bf16 *__restrict__ gm, float *__restrict__ db)
{
    const int c = blockIdx.y;
    float s = 0.f;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x)
        s += bf2f(gm[p * C + c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[c], red[0]);
}

// output 1x1x1 convolution (C -> 1), FP32; also writes y * out_scale into the
// DA-layout double field (interior points) when yd != nullptr
__global__ void k_out(Dims d, Grid g, int C, const bf16 *__restrict__ Y, const float *__restrict__ w, const float *__restrict__ b,
                      float *__restrict__ out, const double *__restrict__ oscale, double *__restrict__ yd)
{
    long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= d.np()) return;
    float s = b[0];
    for (int c = 0; c < C; c++) s += bf2f(Y[p * C + c]) * w[c];
    out[p] = s;
    if (yd) {
        const int i = p % d.ni, j = (p / d.ni) % d.nj, k = p / ((long long)d.ni * d.nj);
        yd[((long long)(k + 1) * g.my + j + 1) * g.mx + i + 1] = (double)s * oscale[0];
    }
}

// MSE loss on the normalised target t = phi * tscale; g_out = 2 (out - t) / N;
// per-block partial loss sums
__global__ void k_loss(Dims d, Grid g, const float *__restrict__ out, const double *__restrict__ phi,
                       const double *__restrict__ tscale, float *__restrict__ gout, double *__restrict__ part)
{
    double s = 0;
    const long long n = d.np();
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const int i = p % d.ni, j = (p / d.ni) % d.nj, k = p / ((long long)d.ni * d.nj);
        const double t = phi[((long long)(k + 1) * g.my + j + 1) * g.mx + i + 1] * tscale[0];
        const double e = (double)out[p] - t;
        s += e * e;
        gout[p] = (float)(2.0 * e / (double)n);
    }
    s = block_reduce<OpSum>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// backward of the output convolution: g_y[p][c] = g_out[p] w[c]; dw, db by partials
__global__ void k_out_back(Dims d, int C, const bf16 *__restrict__ Y, const float *__restrict__ gout, const float *__restrict__ w,
                           float *__restrict__ gy, float *__restrict__ dw, float *__restrict__ db)
{
    float sw[64];
    for (int c = 0; c < C; c++) sw[c] = 0.f;
    float sb = 0.f;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < d.np(); p += (long long)gridDim.x * blockDim.x) {
        const float go = gout[p];
        for (int c = 0; c < C; c++) {
            gy[p * C + c] = go * w[c];
            sw[c] += go * bf2f(Y[p * C + c]);
        }
        sb += go;
    }
    __shared__ float red[256];
    for (int c = 0; c <= C; c++) {
        red[threadIdx.x] = c < C ? sw[c] : sb;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(c < C ? &dw[c] : &db[0], red[0]);
        __syncthreads();
    }
}

__global__ void k_adam(long long n, float lr, float b1, float b2, float eps, float c1, float c2, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float *__restrict__ w)
{
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gg = g[q];
    const float mm = b1 * m[q] + (1.f - b1) * gg;
    const float vv = b2 * v[q] + (1.f - b2) * gg * gg;
    m[q] = mm;
    v[q] = vv;
    w[q] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// rms of the interior rows of x (deterministic partials)
__global__ void __launch_bounds__(RED_THREADS) k_rms_part(Grid g, const double *__restrict__ x, double *__restrict__ part)
{
    double s = 0;
    const long long ni = g.mx - 2, nj = g.my - 2, nk = g.mz - 2, n = ni * nj * nk;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long i = p % ni, j = (p / ni) % nj, k = p / (ni * nj);
        const double v = x[((k + 1) * g.my + j + 1) * g.mx + i + 1];
        s += v * v;
    }
    s = block_reduce<OpSum>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sc[0] = a / sqrt(sum / n) (or 0), sc[1] = sqrt(sum / n) * b
__global__ void k_scales(const double *__restrict__ sum, double n, double a, double b, double *__restrict__ sc)
{
    const double r = sqrt(sum[0] / n);
    sc[0] = r > 0 ? a / r : 0.0;
    sc[1] = r * b;
}

}  // namespace vfs

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace vfs {

struct Conv {
    int cin = 0, cout = 0, level = 0;
    long long nw = 0;                 // 27 * cin * cout
    float *w = nullptr, *b = nullptr;  // FP32 masters
    float *gw = nullptr, *gb = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    bf16x4 *wf = nullptr, *wt = nullptr;  // packed forward / transposed fragments
};

}  // namespace vfs

struct vfs_unet {
    vfs_ctx *ctx = nullptr;
    int levels = 0, base = 0;
    std::vector<vfs::Dims> dims;
    std::vector<vfs::Conv> enc, dec;  // enc[l]: level l, dec[l]: level l (l < levels-1)
    float *wo = nullptr, *bo = nullptr, *gwo = nullptr, *gbo = nullptr, *mwo = nullptr, *vwo = nullptr, *mbo = nullptr,
          *vbo = nullptr;
    // activations (BF16) and gradients (FP32)
    vfs::bf16 *x0 = nullptr;
    std::vector<vfs::bf16 *> e, pool, cat, d;  // e[l], pool[l] (input of enc l, l >= 1), cat[l], d[l]
    std::vector<float *> ge, gpool, gcat, gd;
    vfs::bf16 *gm = nullptr;  // masked BF16 gradient scratch
    float *out = nullptr, *gout = nullptr;
    double *sc = nullptr;      // [0] input scale, [1] output scale, [2..3] target scales
    double *part = nullptr;
    double out_scale = 1.0;    // phi / rms(b) of the training data
    int step = 0;
    std::vector<void *> allocs;
};

namespace vfs {

static int ualloc(vfs_unet *u, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return check_hip(e, "hipMalloc (unet)");
    u->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes ? bytes : 16);
    return e == hipSuccess ? VFS_OK : check_hip(e, "hipMemset (unet)");
}

template <int NT, bool RELU, bool OUT_F32>
static void conv_launch(hipStream_t st, const Dims &d, int cin, const bf16 *X, const bf16x4 *Wf, const float *bias, bool flip,
                        void *Y)
{
    const unsigned nb = (unsigned)(((d.ni + 63) / 64) * (long long)d.nj * d.nk);
    hipLaunchKernelGGL((k_conv3<NT, RELU, OUT_F32>), dim3(nb), dim3(CV_TPB), 0, st, d, cin, X, Wf, bias, flip, Y);
}

// forward convolution with ReLU (BF16 out)
static int conv_fwd(hipStream_t st, const Dims &d, const Conv &c, const bf16 *X, bf16 *Y)
{
    switch (c.cout / 16) {
    case 1: conv_launch<1, true, false>(st, d, c.cin, X, c.wf, c.b, false, Y); break;
    case 2: conv_launch<2, true, false>(st, d, c.cin, X, c.wf, c.b, false, Y); break;
    case 4: conv_launch<4, true, false>(st, d, c.cin, X, c.wf, c.b, false, Y); break;
    case 8: conv_launch<8, true, false>(st, d, c.cin, X, c.wf, c.b, false, Y); break;
    default: return set_error(VFS_EINVAL, "unet: unsupported output channel count");
    }
    VFS_LAUNCH_CHECK("unet conv");
    return VFS_OK;
}

// data gradient dX (FP32) from the masked output gradient G (BF16)
static int conv_bwd_data(hipStream_t st, const Dims &d, const Conv &c, const bf16 *G, float *dX)
{
    switch (c.cin / 16) {
    case 1: conv_launch<1, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 2: conv_launch<2, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 3: conv_launch<3, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 4: conv_launch<4, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 6: conv_launch<6, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 8: conv_launch<8, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    case 12: conv_launch<12, false, true>(st, d, c.cout, G, c.wt, nullptr, true, dX); break;
    default: return set_error(VFS_EINVAL, "unet: unsupported input channel count");
    }
    VFS_LAUNCH_CHECK("unet conv data gradient");
    return VFS_OK;
}

static int conv_bwd_weight(hipStream_t st, const Dims &d, const Conv &c, const bf16 *X, const bf16 *G)
{
    const long long rows = (long long)((d.ni + 63) / 64) * d.nj * d.nk;
    dim3 grid((unsigned)((rows + WG_ROWS - 1) / WG_ROWS), (unsigned)(27 * (c.cin / 16) * (c.cout / 16)));
    hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(64), 0, st, d, c.cin, c.cout, X, G, c.gw);
    hipLaunchKernelGGL(k_bias_grad, dim3(64, c.cout), dim3(256), 0, st, d.np(), c.cout, G, c.gb);
    VFS_LAUNCH_CHECK("unet conv weight gradient");
    return VFS_OK;
}

static int pack(hipStream_t st, Conv &c)
{
    const long long n = 27LL * (c.cin / 16) * (c.cout / 16) * 64;
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(n, 256)), dim3(256), 0, st, c.cin, c.cout, 0, c.w, c.wf);
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(n, 256)), dim3(256), 0, st, c.cin, c.cout, 1, c.w, c.wt);
    VFS_LAUNCH_CHECK("unet pack");
    return VFS_OK;
}

static int conv_alloc(vfs_unet *u, Conv &c, int cin, int cout, int level, int fan_cin, std::mt19937_64 &rng)
{
    c.cin = cin;
    c.cout = cout;
    c.level = level;
    c.nw = 27LL * cin * cout;
    int rc = 0;
    float **fs[] = {&c.w, &c.gw, &c.mw, &c.vw};
    for (float **f : fs) rc = rc ? rc : ualloc(u, (void **)f, c.nw * 4);
    float **bs[] = {&c.b, &c.gb, &c.mb, &c.vb};
    for (float **f : bs) rc = rc ? rc : ualloc(u, (void **)f, cout * 4);
    const long long nf = 27LL * (cin / 16) * (cout / 16) * 64;
    rc = rc ? rc : ualloc(u, (void **)&c.wf, nf * sizeof(bf16x4));
    rc = rc ? rc : ualloc(u, (void **)&c.wt, nf * sizeof(bf16x4));
    if (rc) return rc;
    // He initialisation over the real fan-in (the first layer has one live input channel)
    const int fan = 27 * fan_cin;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan));
    std::vector<float> w(c.nw);
    for (auto &x : w) x = nd(rng);
    VFS_HIP(hipMemcpy(c.w, w.data(), c.nw * 4, hipMemcpyHostToDevice));
    return VFS_OK;
}

static int forward(vfs_unet *u, const double *b_dev, double *y_dev)
{
    vfs_ctx *c = u->ctx;
    hipStream_t st = c->stream;
    const int L = u->levels;
    // input scale 1 / rms(b) and output scale rms(b) * out_scale
    hipLaunchKernelGGL(k_rms_part, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, c->g, b_dev, u->part);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, u->part, RED_BLOCKS, u->part + RED_BLOCKS);
    hipLaunchKernelGGL(k_scales, dim3(1), dim3(1), 0, st, u->part + RED_BLOCKS, (double)u->dims[0].np(), 1.0, u->out_scale, u->sc);
    const Dims &d0 = u->dims[0];
    hipLaunchKernelGGL(k_in, dim3(blocks_for(d0.np(), 256)), dim3(256), 0, st, d0, c->g, b_dev, u->sc, u->x0);
    VFS_LAUNCH_CHECK("unet input");
    int rc = conv_fwd(st, d0, u->enc[0], u->x0, u->e[0]);
    for (int l = 1; l < L && !rc; l++) {
        const int C = u->enc[l - 1].cout;
        hipLaunchKernelGGL(k_pool, dim3(blocks_for(u->dims[l].np() * C, 256)), dim3(256), 0, st, u->dims[l], C, u->e[l - 1],
                           u->pool[l]);
        rc = conv_fwd(st, u->dims[l], u->enc[l], u->pool[l], u->e[l]);
    }
    for (int l = L - 2; l >= 0 && !rc; l--) {
        const bf16 *coarse = l == L - 2 ? u->e[L - 1] : u->d[l + 1];
        const int Cu = l == L - 2 ? u->enc[L - 1].cout : u->dec[l + 1].cout, Cs = u->enc[l].cout;
        hipLaunchKernelGGL(k_up_cat, dim3(blocks_for(u->dims[l].np() * (Cu + Cs), 256)), dim3(256), 0, st, u->dims[l], Cu, Cs,
                           coarse, u->e[l], u->cat[l]);
        rc = conv_fwd(st, u->dims[l], u->dec[l], u->cat[l], u->d[l]);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_out, dim3(blocks_for(d0.np(), 256)), dim3(256), 0, st, d0, c->g, u->dec[0].cout, u->d[0], u->wo, u->bo,
                       u->out, u->sc + 1, y_dev);
    VFS_LAUNCH_CHECK("unet output");
    return VFS_OK;
}

static int backward(vfs_unet *u)
{
    vfs_ctx *c = u->ctx;
    hipStream_t st = c->stream;
    const int L = u->levels;
    const Dims &d0 = u->dims[0];
    auto zero = [&](float *p, long long n) { return hipMemsetAsync(p, 0, n * 4, st); };
    for (auto &cv : u->enc) { zero(cv.gw, cv.nw); zero(cv.gb, cv.cout); }
    for (auto &cv : u->dec) { zero(cv.gw, cv.nw); zero(cv.gb, cv.cout); }
    zero(u->gwo, u->dec[0].cout);
    zero(u->gbo, 1);
    for (int l = 0; l < L; l++) zero(u->ge[l], u->dims[l].np() * u->enc[l].cout);
    hipLaunchKernelGGL(k_out_back, dim3(256), dim3(256), 0, st, d0, u->dec[0].cout, u->d[0], u->gout, u->wo, u->gd[0], u->gwo,
                       u->gbo);
    int rc = 0;
    for (int l = 0; l <= L - 2 && !rc; l++) {
        Conv &cv = u->dec[l];
        const Dims &d = u->dims[l];
        const long long n = d.np() * cv.cout;
        hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, u->d[l], u->gd[l], u->gm);
        rc = conv_bwd_weight(st, d, cv, u->cat[l], u->gm);
        if (!rc) rc = conv_bwd_data(st, d, cv, u->gm, u->gcat[l]);
        if (rc) break;
        const int Cu = cv.cin - u->enc[l].cout, Cs = u->enc[l].cout;
        hipLaunchKernelGGL(k_cat_back_skip, dim3(blocks_for(d.np() * Cs, 256)), dim3(256), 0, st, d, Cu, Cs, u->gcat[l], u->ge[l]);
        float *gcoarse = l == L - 2 ? u->ge[L - 1] : u->gd[l + 1];
        hipLaunchKernelGGL(k_up_cat_back_coarse, dim3(blocks_for(u->dims[l + 1].np() * Cu, 256)), dim3(256), 0, st, u->dims[l + 1],
                           Cu, Cs, u->gcat[l], gcoarse);
        VFS_LAUNCH_CHECK("unet decoder backward");
    }
    for (int l = L - 1; l >= 0 && !rc; l--) {
        Conv &cv = u->enc[l];
        const Dims &d = u->dims[l];
        const long long n = d.np() * cv.cout;
        hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, u->e[l], u->ge[l], u->gm);
        const bf16 *X = l == 0 ? u->x0 : u->pool[l];
        rc = conv_bwd_weight(st, d, cv, X, u->gm);
        if (rc || l == 0) break;
        rc = conv_bwd_data(st, d, cv, u->gm, u->gpool[l]);
        if (rc) break;
        hipLaunchKernelGGL(k_pool_back, dim3(blocks_for(u->dims[l - 1].np() * cv.cin, 256)), dim3(256), 0, st, u->dims[l - 1],
                           cv.cin, u->gpool[l], u->ge[l - 1]);
        VFS_LAUNCH_CHECK("unet encoder backward");
    }
    return rc;
}

static int loss_of(vfs_unet *u, const double *phi_dev, double *loss)
{
    vfs_ctx *c = u->ctx;
    hipStream_t st = c->stream;
    const Dims &d0 = u->dims[0];
    // target scale 1 / (rms(b) * out_scale) = sc[0] / out_scale
    hipLaunchKernelGGL(k_scales, dim3(1), dim3(1), 0, st, u->part + RED_BLOCKS, (double)d0.np(), 1.0 / u->out_scale, 0.0, u->sc + 2);
    hipLaunchKernelGGL(k_loss, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, d0, c->g, u->out, phi_dev, u->sc + 2, u->gout, u->part);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, u->part, RED_BLOCKS, u->part + RED_BLOCKS + 1);
    VFS_LAUNCH_CHECK("unet loss");
    double h[2];
    VFS_HIP(hipMemcpyAsync(h, u->part + RED_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    if (loss) *loss = h[1] / (double)d0.np();
    return VFS_OK;
}

static void for_each_param(vfs_unet *u, const std::function<void(float *, float *, float *, float *, long long)> &f)
{
    for (auto *v : {&u->enc, &u->dec})
        for (auto &cv : *v) {
            f(cv.w, cv.gw, cv.mw, cv.vw, cv.nw);
            f(cv.b, cv.gb, cv.mb, cv.vb, cv.cout);
        }
    f(u->wo, u->gwo, u->mwo, u->vwo, u->dec[0].cout);
    f(u->bo, u->gbo, u->mbo, u->vbo, 1);
}

static int repack(vfs_unet *u)
{
    int rc = 0;
    for (auto &cv : u->enc) rc = rc ? rc : pack(u->ctx->stream, cv);
    for (auto &cv : u->dec) rc = rc ? rc : pack(u->ctx->stream, cv);
    return rc;
}

}  // namespace vfs

using namespace vfs;

extern "C" {

int vfs_unet_create(vfs_ctx *c, int base, int levels, unsigned long long seed, vfs_unet **out)
{
    if (!c || !out || base <= 0 || base % 16 || base > 64 || levels < 2 || levels > 5)
        return set_error(VFS_EINVAL, "vfs_unet_create: base must be 16, 32, 48 or 64 and levels in [2, 5]");
    if (base << (levels - 1) > 128) return set_error(VFS_EINVAL, "vfs_unet_create: at most 128 channels per level");
    const int ni = c->g.mx - 2, nj = c->g.my - 2, nk = c->g.mz - 2, f = 1 << (levels - 1);
    if (ni % f || nj % f || nk % f)
        return set_error(VFS_EINVAL, "vfs_unet_create: interior cell counts must be divisible by 2^(levels-1)");
    vfs_unet *u = new vfs_unet();
    u->ctx = c;
    u->levels = levels;
    u->base = base;
    for (int l = 0; l < levels; l++) u->dims.push_back(Dims{ni >> l, nj >> l, nk >> l});
    std::mt19937_64 rng(seed);
    int rc = 0;
    u->enc.resize(levels);
    u->dec.resize(levels - 1);
    for (int l = 0; l < levels && !rc; l++)
        rc = conv_alloc(u, u->enc[l], l == 0 ? 16 : base << (l - 1), base << l, l, l == 0 ? 1 : base << (l - 1), rng);
    for (int l = levels - 2; l >= 0 && !rc; l--) {
        const int cu = l == levels - 2 ? (base << (levels - 1)) : (base << (l + 1));
        rc = conv_alloc(u, u->dec[l], cu + (base << l), base << l, l, cu + (base << l), rng);
    }
    float **os[] = {&u->wo, &u->gwo, &u->mwo, &u->vwo};
    for (float **p : os) rc = rc ? rc : ualloc(u, (void **)p, base * 4);
    float **ob[] = {&u->bo, &u->gbo, &u->mbo, &u->vbo};
    for (float **p : ob) rc = rc ? rc : ualloc(u, (void **)p, 4);
    if (!rc) {
        std::normal_distribution<float> nd(0.f, std::sqrt(1.f / (float)base));
        std::vector<float> w(base);
        for (auto &x : w) x = nd(rng);
        rc = hipMemcpy(u->wo, w.data(), base * 4, hipMemcpyHostToDevice) == hipSuccess ? 0 : check_hip(hipErrorUnknown, "unet init");
    }
    u->e.assign(levels, nullptr); u->pool.assign(levels, nullptr); u->cat.assign(levels, nullptr); u->d.assign(levels, nullptr);
    u->ge.assign(levels, nullptr); u->gpool.assign(levels, nullptr); u->gcat.assign(levels, nullptr); u->gd.assign(levels, nullptr);
    long long gmax = 0;
    for (int l = 0; l < levels && !rc; l++) {
        const long long np = u->dims[l].np();
        const int C = u->enc[l].cout;
        rc = rc ? rc : ualloc(u, (void **)&u->e[l], np * C * 2);
        rc = rc ? rc : ualloc(u, (void **)&u->ge[l], np * C * 4);
        gmax = std::max(gmax, np * C);
        if (l >= 1) {
            rc = rc ? rc : ualloc(u, (void **)&u->pool[l], np * u->enc[l].cin * 2);
            rc = rc ? rc : ualloc(u, (void **)&u->gpool[l], np * u->enc[l].cin * 4);
        }
        if (l <= levels - 2) {
            rc = rc ? rc : ualloc(u, (void **)&u->cat[l], np * u->dec[l].cin * 2);
            rc = rc ? rc : ualloc(u, (void **)&u->gcat[l], np * u->dec[l].cin * 4);
            rc = rc ? rc : ualloc(u, (void **)&u->d[l], np * u->dec[l].cout * 2);
            rc = rc ? rc : ualloc(u, (void **)&u->gd[l], np * u->dec[l].cout * 4);
            gmax = std::max(gmax, np * u->dec[l].cout);
        }
    }
    const long long np0 = u->dims[0].np();
    rc = rc ? rc : ualloc(u, (void **)&u->x0, np0 * 16 * 2);
    rc = rc ? rc : ualloc(u, (void **)&u->gm, gmax * 2);
    rc = rc ? rc : ualloc(u, (void **)&u->out, np0 * 4);
    rc = rc ? rc : ualloc(u, (void **)&u->gout, np0 * 4);
    rc = rc ? rc : ualloc(u, (void **)&u->sc, 8 * sizeof(double));
    rc = rc ? rc : ualloc(u, (void **)&u->part, (RED_BLOCKS + 8) * sizeof(double));
    if (!rc) rc = repack(u);
    if (rc) { vfs_unet_destroy(u); return rc; }
    *out = u;
    return VFS_OK;
}

int vfs_unet_destroy(vfs_unet *u)
{
    if (!u) return VFS_OK;
    hipStreamSynchronize(u->ctx->stream);
    for (void *p : u->allocs) hipFree(p);
    delete u;
    return VFS_OK;
}

int vfs_unet_set_output_scale(vfs_unet *u, double s)
{
    if (!u || !(s > 0)) return set_error(VFS_EINVAL, "vfs_unet_set_output_scale: positive scale required");
    u->out_scale = s;
    return VFS_OK;
}

int vfs_unet_forward(vfs_unet *u, const double *b_dev, double *y_dev)
{
    if (!u || !b_dev || !y_dev) return set_error(VFS_EINVAL, "vfs_unet_forward: null argument");
    VFS_HIP(hipMemsetAsync(y_dev, 0, (size_t)u->ctx->g.n * sizeof(double), u->ctx->stream));
    return forward(u, b_dev, y_dev);
}

int vfs_unet_gradients(vfs_unet *u, const double *b_dev, const double *phi_dev, double *loss)
{
    if (!u || !b_dev || !phi_dev) return set_error(VFS_EINVAL, "vfs_unet_gradients: null argument");
    int rc = forward(u, b_dev, nullptr);
    if (!rc) rc = loss_of(u, phi_dev, loss);
    if (!rc) rc = backward(u);
    return rc;
}

int vfs_unet_train_step(vfs_unet *u, const double *b_dev, const double *phi_dev, double lr, double *loss)
{
    int rc = vfs_unet_gradients(u, b_dev, phi_dev, loss);
    if (rc) return rc;
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    hipStream_t st = u->ctx->stream;
    for_each_param(u, [&](float *w, float *g, float *m, float *v, long long n) {
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, (float)lr, b1, b2, eps, c1, c2, g, m, v, w);
    });
    VFS_LAUNCH_CHECK("unet adam");
    return repack(u);
}

long long vfs_unet_param_count(vfs_unet *u)
{
    if (!u) return 0;
    long long n = 0;
    for_each_param(u, [&](float *, float *, float *, float *, long long k) { n += k; });
    return n;
}

// which: 0 = weights (FP32 masters), 1 = gradients of the last vfs_unet_gradients
int vfs_unet_get_params(vfs_unet *u, int which, float *host)
{
    if (!u || !host) return set_error(VFS_EINVAL, "vfs_unet_get_params: null argument");
    VFS_HIP(hipStreamSynchronize(u->ctx->stream));
    long long off = 0;
    int rc = 0;
    for_each_param(u, [&](float *w, float *g, float *, float *, long long n) {
        if (!rc && hipMemcpy(host + off, which ? g : w, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_error(VFS_EHIP, "vfs_unet_get_params: copy failed");
        off += n;
    });
    return rc;
}

int vfs_unet_set_params(vfs_unet *u, const float *host)
{
    if (!u || !host) return set_error(VFS_EINVAL, "vfs_unet_set_params: null argument");
    VFS_HIP(hipStreamSynchronize(u->ctx->stream));
    long long off = 0;
    int rc = 0;
    for_each_param(u, [&](float *w, float *, float *, float *, long long n) {
        if (!rc && hipMemcpy(w, host + off, n * 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = set_error(VFS_EHIP, "vfs_unet_set_params: copy failed");
        off += n;
    });
    return rc ? rc : repack(u);
}

}  // extern "C"
