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
// cores, FP32 master weights, gradients and Adam state.  Activations are
// channels-last ([k][j][i][c], i fastest among the points) so that one lane's
// eight consecutive channels are one 16-byte load.
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x32_bf16: M = points (one
// MFMA tile = 16 consecutive points along i), N = output channels (16 per
// tile), K = taps x input channels.  A workgroup owns a 16 x 4*MT x 1 tile of
// points and stages its 18 x (4*MT+2) x 3 halo in LDS, CC input channels at a
// time; within a chunk K runs over (tap, channel) flattened, 32 per MFMA, so
// a lane's eight K values are eight consecutive channels of one tap.  The data
// gradient is the same kernel on the output gradient with the weights
// transposed and the taps mirrored (both folded into the packing); the weight
// gradient is a second MFMA kernel with K = points, reduced over fixed
// point ranges in a fixed order (no atomics: training is deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"

namespace vfs {

typedef uint16_t bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));     // eight BF16 as stored
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));   // MFMA operand
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

constexpr int CV_TPB = 256;  // 4 waves

// Channel chunk staged in LDS per pass: 32 when the K-channel count allows, else 16.
__host__ __device__ constexpr int conv_cc(int kc) { return kc % 32 == 0 ? 32 : 16; }
// MFMA K steps per chunk: 27 taps x CC channels, 32 per step (the last one zero padded).
__host__ __device__ constexpr int conv_ks(int cc) { return (27 * cc + 31) / 32; }
// M tiles per wave: fewer for wide outputs (the accumulators are MT x NT x 4 VGPRs).
__host__ __device__ constexpr int conv_mt(int nt) { return nt <= 2 ? 4 : nt <= 4 ? 2 : 1; }
// LDS bytes of the halo tile: 18 x (4 MT + 2) x 3 points of CC channels, each
// point padded by 16 bytes so that 16 lanes reading 16 consecutive points hit
// distinct banks.
__host__ __device__ constexpr int conv_lds(int mt, int cc) { return 18 * (4 * mt + 2) * 3 * (2 * cc + 16); }

// Y[p][co] = act(sum_{tap, ci} X[p + tap][ci] W[tap][ci][co] + bias[co]) for
// the points of one 16 x 4*MT x 1 tile.  Wf holds the B fragments in the
// order the K loop consumes them: Wf[((chunk * KS + s) * NT + t) * 64 + lane]
// = the eight weights of flattened K = 32 s + 8 (lane / 16) + 0..7 (tap = K /
// CC, channel = chunk * CC + K % CC) and output channel 16 t + lane % 16.
template <int NT, int MT, int CC, bool RELU, bool OUT_F32>
__global__ void __launch_bounds__(CV_TPB) k_conv3t(Dims d, int cin, const bf16 *__restrict__ X, const bf16x8 *__restrict__ Wf,
                                                   const float *__restrict__ bias, void *__restrict__ Yv)
{
    constexpr int TJ = 4 * MT, HI = 18, HJ = TJ + 2, HP = HI * HJ * 3;
    constexpr int PS = 2 * CC + 16;  // bytes per halo point
    constexpr int KS = conv_ks(CC);
    constexpr int U = CC / 8;        // 16-byte units per halo point
    extern __shared__ __align__(16) unsigned char lds[];
    const int cout = 16 * NT;
    const int nbi = (d.ni + 15) / 16, nbj = (d.nj + TJ - 1) / TJ;
    const int bi = blockIdx.x % nbi, bj = (blockIdx.x / nbi) % nbj, k = blockIdx.x / (nbi * nbj);
    const int i0 = bi * 16, j0 = bj * TJ;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int r = 0; r < MT; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this lane's halo point for tap (0, 0, 0) in each of its rows
    int base[MT];
#pragma unroll
    for (int r = 0; r < MT; r++) base[r] = (HJ + wave * MT + r + 1) * HI + m + 1;
    const int nch = cin / CC;
    for (int cc = 0; cc < nch; cc++) {
        if (cc) __syncthreads();  // every wave is done with the previous chunk
        for (int u = threadIdx.x; u < HP * U; u += CV_TPB) {
            const int hp = u / U, c8 = u - hp * U;
            const int hi = hp % HI, hj = (hp / HI) % HJ, hk = hp / (HI * HJ);
            const int ii = i0 - 1 + hi, jj = j0 - 1 + hj, kk = k - 1 + hk;
            s16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
            if (ii >= 0 && ii < d.ni && jj >= 0 && jj < d.nj && kk >= 0 && kk < d.nk)
                v = *reinterpret_cast<const s16x8 *>(X + (((long long)kk * d.nj + jj) * d.ni + ii) * cin + cc * CC + 8 * c8);
            *reinterpret_cast<s16x8 *>(lds + hp * PS + 16 * c8) = v;
        }
        __syncthreads();
        const bf16x8 *wc = Wf + (long long)cc * KS * NT * 64 + lane;
#pragma unroll 2
        for (int s = 0; s < KS; s++) {
            const int kf = 32 * s + 8 * g;
            const int tap = kf / CC, ch = kf - tap * CC;
            bf16x8 b[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) b[t] = wc[(s * NT + t) * 64];
            bf16x8 a[MT];
            if (tap < 27) {
                const int dk = tap / 9 - 1, dj = (tap / 3) % 3 - 1, di = tap % 3 - 1;
                const int off = (dk * HJ + dj) * HI + di;
#pragma unroll
                for (int r = 0; r < MT; r++)
                    a[r] = *reinterpret_cast<const bf16x8 *>(lds + (base[r] + off) * PS + 2 * ch);
            } else {
#pragma unroll
                for (int r = 0; r < MT; r++) a[r] = bf16x8{};
            }
#pragma unroll
            for (int r = 0; r < MT; r++)
#pragma unroll
                for (int t = 0; t < NT; t++) acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[t], acc[r][t], 0, 0, 0);
        }
    }
    // lane holds rows 4 g + q (points along i) of column m (channel) of each tile
#pragma unroll
    for (int r = 0; r < MT; r++) {
        const int jp = j0 + wave * MT + r;
        if (jp >= d.nj) continue;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = 16 * t + m;
            const float bb = bias ? bias[co] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int ip = i0 + 4 * g + q;
                if (ip >= d.ni) continue;
                float v = acc[r][t][q] + bb;
                if (RELU) v = v > 0.f ? v : 0.f;
                const long long o = (((long long)k * d.nj + jp) * d.ni + ip) * cout + co;
                if (OUT_F32) reinterpret_cast<float *>(Yv)[o] = v;
                else reinterpret_cast<bf16 *>(Yv)[o] = f2bf(v);
            }
        }
    }
}

// Partial weight gradients: block (x, y) sums X[p + tap][ci] G[p][co] over
// the rows x, x + P, x + 2P, ... (P = gridDim.x; a row is a 64-point run along
// i) for its (tap, 16-channel ci tile, 16-channel co tile) and writes the tile
// into slice x of `part` (each slice 27 cin cout floats); k_sum_partials then
// adds the slices in a fixed order, so the gradient does not depend on timing.
__global__ void __launch_bounds__(64) k_conv3_wgrad(Dims d, int cin, int cout, const bf16 *__restrict__ X,
                                                    const bf16 *__restrict__ G, float *__restrict__ part)
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
    for (long long row = blockIdx.x; row < nrows; row += gridDim.x) {
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
    // acc[r] = dW[ci = 4 kq + r][co = m] of this tile
    float *slice = part + (long long)blockIdx.x * 27 * cin * cout;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int ci = cb * 16 + 4 * kq + r, co = t * 16 + m;
        slice[((long long)tap * cin + ci) * cout + co] = acc[r];
    }
}

// out[i] = sum over p = 0 .. P-1, in order, of part[p * stride + i], i < n
__global__ void k_sum_partials(const float *__restrict__ part, int P, long long stride, long long n, float *__restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int p = 0; p < P; p++) s += part[p * stride + i];
    out[i] = s;
}

// The same convolution marching along k, for layers whose three halo planes
// of all KC input channels and whole packed weight set fit in LDS together
// (conv_m_lds): a block owns a 16 x 4*MT column of points over a run of kr
// planes.  Three halo planes sit in an LDS ring; while the matrix cores work
// on plane k, each thread's share of plane k + 2 is already in flight into
// registers, and it replaces plane k - 1 in the ring afterwards.  Each input
// plane is read from memory once per run instead of three times, and the
// weights once per block.  Wf is packed as one chunk of KC channels
// (k_pack with cc_all): Wf[(s * NT + t) * 64 + lane].
// 16-output layers run eight waves of two M tiles (the same 16 x 16 tile as
// four waves of four: twice the waves per SIMD to hide the LDS latency of
// their one-MFMA-per-A-fragment inner loop); wider layers four waves.
__host__ __device__ constexpr int conv_m_waves(int nt) { return nt == 1 ? 8 : 4; }
__host__ __device__ constexpr int conv_m_mt(int nt) { return nt == 1 ? 2 : conv_mt(nt); }
__host__ __device__ constexpr int conv_m_tj(int nt) { return conv_m_waves(nt) * conv_m_mt(nt); }
__host__ __device__ constexpr int conv_m_ring(int tj, int kc) { return 3 * 18 * (tj + 2) * (2 * kc + 16); }
__host__ __device__ constexpr int conv_m_lds(int nt, int kc)
{
    return conv_m_ring(conv_m_tj(nt), kc) + conv_ks(kc) * nt * 64 * 16;
}

// UPCAT: the decoder's concatenation is read in place: input channels
// [0, cu) are the half-resolution tensor Xc upsampled (nearest), channels
// [cu, KC) the skip tensor X (KC - cu channels), so the concatenated tensor
// is never written.
template <int NT, int MT, int KC, bool RELU, bool OUT_F32, bool UPCAT, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_conv3m(Dims d, int kr, const bf16 *__restrict__ X, const bf16 *__restrict__ Xc,
                                                       int cu, const s16x8 *__restrict__ Wf, const float *__restrict__ bias,
                                                       void *__restrict__ Yv)
{
    constexpr int TPB = 64 * WAVES;
    constexpr int TJ = WAVES * MT, HI = 18, HJ = TJ + 2, PP = HI * HJ;
    constexpr int PS = 2 * KC + 16;          // bytes per halo point (16 lanes, 16 points: distinct banks)
    constexpr int U = KC / 8;                // 16-byte units per halo point
    constexpr int KS = conv_ks(KC);
    constexpr int NPF = (PP * U + TPB - 1) / TPB;  // halo units per thread and plane
    constexpr int PLANE = PP * PS, RING = 3 * PLANE;
    extern __shared__ __align__(16) unsigned char lds[];
    const int cout = 16 * NT;
    const int nbi = (d.ni + 15) / 16, nbj = (d.nj + TJ - 1) / TJ;
    const int bi = blockIdx.x % nbi, bj = (blockIdx.x / nbi) % nbj, bk = blockIdx.x / (nbi * nbj);
    const int i0 = bi * 16, j0 = bj * TJ, k0 = bk * kr;
    const int k1 = k0 + kr < d.nk ? k0 + kr : d.nk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int cs = UPCAT ? KC - cu : KC;                // channels read from X
    const long long xplane = (long long)d.ni * d.nj * cs;
    const long long cplane = UPCAT ? (long long)(d.ni / 2) * (d.nj / 2) * cu : 0;
    // the layer's weights, once per block
    s16x8 *Bs = reinterpret_cast<s16x8 *>(lds + RING);
    for (int q = threadIdx.x; q < KS * NT * 64; q += TPB) Bs[q] = Wf[q];
    // this thread's halo units: the same (i, j, channel) offsets on every plane
    int goff[NPF], loff[NPF];
    unsigned coarse = 0;  // bit n: unit n reads the half-resolution tensor
#pragma unroll
    for (int n = 0; n < NPF; n++) {
        const int u = threadIdx.x + n * TPB;
        goff[n] = -1;
        loff[n] = -1;
        if (u < PP * U) {
            const int hp = u / U, c8 = u - hp * U;
            const int hi = hp % HI, hj = hp / HI;
            const int ii = i0 - 1 + hi, jj = j0 - 1 + hj;
            loff[n] = hp * PS + 16 * c8;
            if (ii >= 0 && ii < d.ni && jj >= 0 && jj < d.nj) {
                if (UPCAT && 8 * c8 < cu) {
                    goff[n] = ((jj >> 1) * (d.ni >> 1) + (ii >> 1)) * cu + 8 * c8;
                    coarse |= 1u << n;
                } else {
                    goff[n] = (jj * d.ni + ii) * cs + 8 * c8 - (UPCAT ? cu : 0);
                }
            }
        }
    }
    s16x8 v[NPF];
    auto fetch = [&](int kk) {
        const bool in = kk >= 0 && kk < d.nk;
        const bf16 *xp = X + (long long)kk * xplane;
        const bf16 *cp = UPCAT ? Xc + (long long)(kk >> 1) * cplane : nullptr;
#pragma unroll
        for (int n = 0; n < NPF; n++) {
            v[n] = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (in && goff[n] >= 0)
                v[n] = *reinterpret_cast<const s16x8 *>((UPCAT && (coarse >> n & 1u) ? cp : xp) + goff[n]);
        }
    };
    auto put = [&](int slot) {
#pragma unroll
        for (int n = 0; n < NPF; n++)
            if (loff[n] >= 0) *reinterpret_cast<s16x8 *>(lds + slot * PLANE + loff[n]) = v[n];
    };
    // ring slot of plane kk: (kk - k0 + 1) mod 3
    fetch(k0 - 1);
    put(0);
    fetch(k0);
    put(1);
    fetch(k0 + 1);
    put(2);
    __syncthreads();
    int base[MT];
#pragma unroll
    for (int r = 0; r < MT; r++) base[r] = (wave * MT + r + 1) * HI + m + 1;
    for (int k = k0; k < k1; k++) {
        const bool next = k + 1 < k1;
        if (next) fetch(k + 2);  // in flight during the matrix work
        const int s0 = (k - k0) % 3;  // slot of plane k - 1
        const int sm = s0 * PLANE, sc = (s0 == 2 ? 0 : s0 + 1) * PLANE, sp = (s0 == 0 ? 2 : s0 - 1) * PLANE;
        f32x4 acc[MT][NT];
#pragma unroll
        for (int r = 0; r < MT; r++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int kf = 32 * s + 8 * g;
            const int tap = kf / KC, ch = kf - tap * KC;
            bf16x8 b[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) b[t] = __builtin_bit_cast(bf16x8, Bs[(s * NT + t) * 64 + lane]);
            bf16x8 a[MT];
            if (tap < 27) {
                const int dk = tap / 9, dj = (tap / 3) % 3 - 1, di = tap % 3 - 1;
                const int pl = dk == 0 ? sm : dk == 1 ? sc : sp;
                const int off = dj * HI + di;
#pragma unroll
                for (int r = 0; r < MT; r++)
                    a[r] = *reinterpret_cast<const bf16x8 *>(lds + pl + (base[r] + off) * PS + 2 * ch);
            } else {
#pragma unroll
                for (int r = 0; r < MT; r++) a[r] = bf16x8{};
            }
#pragma unroll
            for (int r = 0; r < MT; r++)
#pragma unroll
                for (int t = 0; t < NT; t++) acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[t], acc[r][t], 0, 0, 0);
        }
        // lane holds rows 4 g + q (points along i) of column m (channel) of each tile
#pragma unroll
        for (int r = 0; r < MT; r++) {
            const int jp = j0 + wave * MT + r;
            if (jp >= d.nj) continue;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int co = 16 * t + m;
                const float bb = bias ? bias[co] : 0.f;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int ip = i0 + 4 * g + q;
                    if (ip >= d.ni) continue;
                    float val = acc[r][t][q] + bb;
                    if (RELU) val = val > 0.f ? val : 0.f;
                    const long long o = (((long long)k * d.nj + jp) * d.ni + ip) * cout + co;
                    if (OUT_F32) reinterpret_cast<float *>(Yv)[o] = val;
                    else reinterpret_cast<bf16 *>(Yv)[o] = f2bf(val);
                }
            }
        }
        if (next) {
            __syncthreads();  // every wave is done with plane k - 1
            put(s0);          // plane k + 2 takes its slot
            __syncthreads();
        }
    }
}

// B-fragment packing (see k_conv3t) of the FP32 master weights W[27][cin][cout]:
// forward (K channels = cin, outputs = cout) or, for the data gradient,
// transposed with the taps mirrored (K channels = cout, outputs = cin,
// W'[tap][co][ci] = W[26 - tap][ci][co]).
// cc = 0 packs in conv_cc(K channels) chunks; cc = K channels packs one chunk
// (the order of k_conv3m).
__global__ void k_pack(int cin, int cout, int transpose, int cc_all, const float *__restrict__ W, s16x8 *__restrict__ Wf)
{
    const int kc = transpose ? cout : cin, oc = transpose ? cin : cout;
    const int CC = cc_all ? kc : conv_cc(kc), KS = conv_ks(CC), NT = oc / 16, nch = kc / CC;
    const long long total = (long long)nch * KS * NT * 64;
    long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int lane = q % 64;
    long long r = q / 64;
    const int t = r % NT;
    r /= NT;
    const int s = r % KS, cc = r / KS;
    const int o = 16 * t + (lane & 15);
    s16x8 v;
    for (int e = 0; e < 8; e++) {
        const int kf = 32 * s + 8 * (lane >> 4) + e;
        const int tap = kf / CC, ch = cc * CC + kf % CC;
        float w = 0.f;
        if (tap < 27) w = transpose ? W[((long long)(26 - tap) * cin + o) * cout + ch] : W[((long long)tap * cin + ch) * cout + o];
        v[e] = (short)f2bf(w);
    }
    Wf[q] = v;
}

static long long packed_count(int kc, int oc, bool cc_all = false)
{
    const int CC = cc_all ? kc : conv_cc(kc);
    return (long long)(kc / CC) * conv_ks(CC) * (oc / 16) * 64;
}

// ---- elementwise layers ------------------------------------------------------
// input channel 0 = x * scale (double -> bf16), channels 1..15 = 0; one
// thread per point, two 16-byte stores
__global__ void k_in(Dims d, Grid g, const double *__restrict__ x, const double *__restrict__ scale, bf16 *__restrict__ X)
{
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned)d.np()) return;
    const unsigned i = p % d.ni, r = p / d.ni, j = r % d.nj, k = r / d.nj;
    const double v = x[((long long)(k + 1) * g.my + j + 1) * g.mx + i + 1] * scale[0];
    s16x8 lo = {0, 0, 0, 0, 0, 0, 0, 0}, hi = lo;
    lo[0] = (short)f2bf((float)v);
    s16x8 *o = reinterpret_cast<s16x8 *>(X + (long long)p * 16);
    o[0] = lo;
    o[1] = hi;
}

// 2x2x2 average pooling, channels-last: one thread per (coarse point, eight
// channels), 16-byte loads of the eight children, summed in child order
__global__ void k_pool(Dims dc, int C, const bf16 *__restrict__ Xf, bf16 *__restrict__ Xc)
{
    const unsigned U = C / 8;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)dc.np() * U) return;
    const unsigned u = q % U, P = q / U;
    const unsigned I = P % dc.ni, R = P / dc.ni, J = R % dc.nj, K = R / dc.nj;
    const unsigned nif = 2 * dc.ni, njf = 2 * dc.nj;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int e = 0; e < 8; e++) {
        const long long pf = ((long long)(2 * K + (e >> 2)) * njf + 2 * J + ((e >> 1) & 1)) * nif + 2 * I + (e & 1);
        const s16x8 v = *reinterpret_cast<const s16x8 *>(Xf + pf * C + 8 * u);
#pragma unroll
        for (int c = 0; c < 8; c++) s[c] += bf2f((bf16)v[c]);
    }
    s16x8 o;
#pragma unroll
    for (int c = 0; c < 8; c++) o[c] = (short)f2bf(s[c] * 0.125f);
    *reinterpret_cast<s16x8 *>(Xc + (long long)q * 8) = o;
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

// concat buffer of the decoder: [upsampled coarse (Cu) | skip (Cs)]; one
// thread per (point, eight channels), 16-byte copies
__global__ void k_up_cat(Dims df, int Cu, int Cs, const bf16 *__restrict__ Yc, const bf16 *__restrict__ S,
                         bf16 *__restrict__ Z)
{
    const unsigned Uu = Cu / 8, U = (Cu + Cs) / 8;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)df.np() * U) return;
    const unsigned u = q % U, p = q / U;
    s16x8 v;
    if (u < Uu) {
        const unsigned i = p % df.ni, r = p / df.ni, j = r % df.nj, k = r / df.nj;
        const long long P = ((long long)(k / 2) * (df.nj / 2) + j / 2) * (df.ni / 2) + i / 2;
        v = *reinterpret_cast<const s16x8 *>(Yc + P * Cu + 8 * u);
    } else {
        v = *reinterpret_cast<const s16x8 *>(S + (long long)p * Cs + 8 * (u - Uu));
    }
    *reinterpret_cast<s16x8 *>(Z + (long long)q * 8) = v;
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

// bias gradient partials: part[x * C + c] = sum of gm[p][c] over block x's grid-stride points
constexpr int BG_BLOCKS = 64;
__global__ void k_bias_grad(long long np, int C, const bf16 *__restrict__ gm, float *__restrict__ part)
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
    if (threadIdx.x == 0) part[blockIdx.x * C + c] = red[0];
}

// output 1x1x1 convolution (C -> 1), FP32; also writes y * out_scale into the
// DA-layout double field (interior points) when yd != nullptr
__global__ void k_out(Dims d, Grid g, int C, const bf16 *__restrict__ Y, const float *__restrict__ w, const float *__restrict__ b,
                      float *__restrict__ out, const double *__restrict__ oscale, double *__restrict__ yd)
{
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned)d.np()) return;
    float s = b[0];
    const s16x8 *y = reinterpret_cast<const s16x8 *>(Y + (long long)p * C);
    for (int u = 0; u < C / 8; u++) {
        const s16x8 v = y[u];
#pragma unroll
        for (int c = 0; c < 8; c++) s += bf2f((bf16)v[c]) * w[8 * u + c];
    }
    out[p] = s;
    if (yd) {
        const unsigned i = p % d.ni, r = p / d.ni, j = r % d.nj, k = r / d.nj;
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

// backward of the output convolution: g_y[p][c] = g_out[p] w[c]; the
// partial sums of dw (C values) and db (1) of block x go to part[x (C + 1) ..]
constexpr int OB_BLOCKS = 256;
__global__ void k_out_back(Dims d, int C, const bf16 *__restrict__ Y, const float *__restrict__ gout, const float *__restrict__ w,
                           float *__restrict__ gy, float *__restrict__ part)
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
        if (threadIdx.x == 0) part[(long long)blockIdx.x * (C + 1) + c] = red[0];
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
    bf16x8 *wf = nullptr, *wt = nullptr;  // packed forward / transposed B fragments
    long long nwf = 0, nwt = 0;           // their counts
    bool march = false;                   // forward by k_conv3m (wf packed as one chunk)
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
    float *wpart = nullptr;    // partial weight / bias gradient slices
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

template <int NT, int CC, bool RELU, bool OUT_F32>
static int conv_launch(hipStream_t st, const Dims &d, int kc, const bf16 *X, const bf16x8 *Wf, const float *bias, void *Y)
{
    constexpr int MT = conv_mt(NT), TJ = 4 * MT, LDS = conv_lds(MT, CC);
    auto kern = k_conv3t<NT, MT, CC, RELU, OUT_F32>;
    static bool attr = false;  // the halo tile may exceed the 64 KiB default of dynamic LDS
    if (!attr) {
        VFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, LDS));
        attr = true;
    }
    const unsigned nb = (unsigned)((long long)((d.ni + 15) / 16) * ((d.nj + TJ - 1) / TJ) * d.nk);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(CV_TPB), LDS, st, d, kc, X, Wf, bias, Y);
    return VFS_OK;
}

template <int NT, bool RELU, bool OUT_F32>
static int conv_launch_cc(hipStream_t st, const Dims &d, int kc, const bf16 *X, const bf16x8 *Wf, const float *bias, void *Y)
{
    return conv_cc(kc) == 32 ? conv_launch<NT, 32, RELU, OUT_F32>(st, d, kc, X, Wf, bias, Y)
                             : conv_launch<NT, 16, RELU, OUT_F32>(st, d, kc, X, Wf, bias, Y);
}

// k_conv3m launch: runs of kr planes, enough blocks to fill the GPU
template <int NT, int KC, bool UPCAT = false>
static int conv_m_launch(hipStream_t st, const Dims &d, const bf16 *X, const bf16x8 *Wf, const float *bias, bf16 *Y,
                         const bf16 *Xc = nullptr, int cu = 0)
{
    constexpr int MT = conv_m_mt(NT), W = conv_m_waves(NT), TJ = conv_m_tj(NT), LDS = conv_m_lds(NT, KC);
    static_assert(LDS <= 160 * 1024, "halo ring and weights exceed the LDS");
    auto kern = k_conv3m<NT, MT, KC, true, false, UPCAT, W>;
    static bool attr = false;
    if (!attr) {
        VFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, LDS));
        attr = true;
    }
    const long long nbp = (long long)((d.ni + 15) / 16) * ((d.nj + TJ - 1) / TJ);
    long long kr = d.nk * nbp / 1024;
    kr = kr < 2 ? 2 : kr > 16 ? 16 : kr;
    if (kr > d.nk) kr = d.nk;
    const unsigned nb = (unsigned)(nbp * ((d.nk + kr - 1) / kr));
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * W), LDS, st, d, (int)kr, X, Xc, cu, reinterpret_cast<const s16x8 *>(Wf),
                       bias, Y);
    return VFS_OK;
}

// layers k_conv3m takes: (output tiles, input channels) with an instance
static bool conv_m_ok(int cin, int cout)
{
    const int nt = cout / 16;
    return (nt == 1 && (cin == 16 || cin == 32 || cin == 48)) || (nt == 2 && (cin == 16 || cin == 32)) ||
           (nt == 4 && cin == 32);
}

// forward convolution with ReLU (BF16 out)
static int conv_fwd(hipStream_t st, const Dims &d, const Conv &c, const bf16 *X, bf16 *Y)
{
    int rc;
    if (c.march) {
        const int key = (c.cout / 16) * 1000 + c.cin;
        switch (key) {
        case 1016: rc = conv_m_launch<1, 16>(st, d, X, c.wf, c.b, Y); break;
        case 1032: rc = conv_m_launch<1, 32>(st, d, X, c.wf, c.b, Y); break;
        case 1048: rc = conv_m_launch<1, 48>(st, d, X, c.wf, c.b, Y); break;
        case 2016: rc = conv_m_launch<2, 16>(st, d, X, c.wf, c.b, Y); break;
        case 2032: rc = conv_m_launch<2, 32>(st, d, X, c.wf, c.b, Y); break;
        case 4032: rc = conv_m_launch<4, 32>(st, d, X, c.wf, c.b, Y); break;
        default: return set_error(VFS_EINVAL, "unet: no k-marching instance for this layer");
        }
        if (rc) return rc;
        VFS_LAUNCH_CHECK("unet conv (k-marching)");
        return VFS_OK;
    }
    switch (c.cout / 16) {
    case 1: rc = conv_launch_cc<1, true, false>(st, d, c.cin, X, c.wf, c.b, Y); break;
    case 2: rc = conv_launch_cc<2, true, false>(st, d, c.cin, X, c.wf, c.b, Y); break;
    case 4: rc = conv_launch_cc<4, true, false>(st, d, c.cin, X, c.wf, c.b, Y); break;
    case 8: rc = conv_launch_cc<8, true, false>(st, d, c.cin, X, c.wf, c.b, Y); break;
    default: return set_error(VFS_EINVAL, "unet: unsupported output channel count");
    }
    if (rc) return rc;
    VFS_LAUNCH_CHECK("unet conv");
    return VFS_OK;
}

// data gradient dX (FP32) from the masked output gradient G (BF16)
static int conv_bwd_data(hipStream_t st, const Dims &d, const Conv &c, const bf16 *G, float *dX)
{
    int rc;
    switch (c.cin / 16) {
    case 1: rc = conv_launch_cc<1, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 2: rc = conv_launch_cc<2, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 3: rc = conv_launch_cc<3, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 4: rc = conv_launch_cc<4, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 6: rc = conv_launch_cc<6, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 8: rc = conv_launch_cc<8, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    case 12: rc = conv_launch_cc<12, false, true>(st, d, c.cout, G, c.wt, nullptr, dX); break;
    default: return set_error(VFS_EINVAL, "unet: unsupported input channel count");
    }
    if (rc) return rc;
    VFS_LAUNCH_CHECK("unet conv data gradient");
    return VFS_OK;
}

constexpr int WG_SLICES = 64;  // partial weight-gradient slices (fixed: the summation order is too)

static int conv_bwd_weight(hipStream_t st, const Dims &d, const Conv &c, const bf16 *X, const bf16 *G, float *wpart)
{
    dim3 grid(WG_SLICES, (unsigned)(27 * (c.cin / 16) * (c.cout / 16)));
    hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(64), 0, st, d, c.cin, c.cout, X, G, wpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(blocks_for(c.nw, 256)), dim3(256), 0, st, wpart, WG_SLICES, c.nw, c.nw, c.gw);
    hipLaunchKernelGGL(k_bias_grad, dim3(BG_BLOCKS, c.cout), dim3(256), 0, st, d.np(), c.cout, G, wpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, wpart, BG_BLOCKS, (long long)c.cout, (long long)c.cout, c.gb);
    VFS_LAUNCH_CHECK("unet conv weight gradient");
    return VFS_OK;
}

static int pack(hipStream_t st, Conv &c)
{
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(c.nwf, 256)), dim3(256), 0, st, c.cin, c.cout, 0, c.march ? 1 : 0, c.w,
                       reinterpret_cast<s16x8 *>(c.wf));
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(c.nwt, 256)), dim3(256), 0, st, c.cin, c.cout, 1, 0, c.w,
                       reinterpret_cast<s16x8 *>(c.wt));
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
    const char *pe = std::getenv("VFS_UNET_KMARCH");  // 0: every layer on k_conv3t
    c.march = !(pe && pe[0] == '0') && conv_m_ok(cin, cout);
    c.nwf = packed_count(cin, cout, c.march);
    c.nwt = packed_count(cout, cin);
    rc = rc ? rc : ualloc(u, (void **)&c.wf, c.nwf * sizeof(bf16x8));
    rc = rc ? rc : ualloc(u, (void **)&c.wt, c.nwt * sizeof(bf16x8));
    if (rc) return rc;
    // He initialisation over the real fan-in (the first layer has one live input channel)
    const int fan = 27 * fan_cin;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan));
    std::vector<float> w(c.nw);
    for (auto &x : w) x = nd(rng);
    VFS_HIP(hipMemcpy(c.w, w.data(), c.nw * 4, hipMemcpyHostToDevice));
    return VFS_OK;
}

// the decoder layers whose concatenation k_conv3m reads in place
static bool upcat_fused(const Conv &c, int cu)
{
    return c.march && c.cout == 16 && c.cin == 48 && cu == 32;
}

// keep_cat: the concatenated decoder inputs are written (the weight
// gradients read them); inference reads them in place where it can
static int forward(vfs_unet *u, const double *b_dev, double *y_dev, bool keep_cat)
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
        hipLaunchKernelGGL(k_pool, dim3(blocks_for(u->dims[l].np() * C / 8, 256)), dim3(256), 0, st, u->dims[l], C, u->e[l - 1],
                           u->pool[l]);
        rc = conv_fwd(st, u->dims[l], u->enc[l], u->pool[l], u->e[l]);
    }
    for (int l = L - 2; l >= 0 && !rc; l--) {
        const bf16 *coarse = l == L - 2 ? u->e[L - 1] : u->d[l + 1];
        const int Cu = l == L - 2 ? u->enc[L - 1].cout : u->dec[l + 1].cout, Cs = u->enc[l].cout;
        if (!keep_cat && upcat_fused(u->dec[l], Cu)) {
            rc = conv_m_launch<1, 48, true>(st, u->dims[l], u->e[l], u->dec[l].wf, u->dec[l].b, u->d[l], coarse, Cu);
            if (!rc) VFS_LAUNCH_CHECK("unet conv (k-marching, concatenation in place)");
            continue;
        }
        hipLaunchKernelGGL(k_up_cat, dim3(blocks_for(u->dims[l].np() * (Cu + Cs) / 8, 256)), dim3(256), 0, st, u->dims[l], Cu, Cs,
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
    for (int l = 0; l < L; l++) zero(u->ge[l], u->dims[l].np() * u->enc[l].cout);
    const int C0 = u->dec[0].cout;
    hipLaunchKernelGGL(k_out_back, dim3(OB_BLOCKS), dim3(256), 0, st, d0, C0, u->d[0], u->gout, u->wo, u->gd[0], u->wpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, u->wpart, OB_BLOCKS, (long long)C0 + 1, (long long)C0, u->gwo);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, u->wpart + C0, OB_BLOCKS, (long long)C0 + 1, 1LL, u->gbo);
    int rc = 0;
    for (int l = 0; l <= L - 2 && !rc; l++) {
        Conv &cv = u->dec[l];
        const Dims &d = u->dims[l];
        const long long n = d.np() * cv.cout;
        hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, u->d[l], u->gd[l], u->gm);
        rc = conv_bwd_weight(st, d, cv, u->cat[l], u->gm, u->wpart);
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
        rc = conv_bwd_weight(st, d, cv, X, u->gm, u->wpart);
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
    // the elementwise layers index (point, eight channels) in 32-bit arithmetic
    if ((long long)ni * nj * nk * (2 * base / 8) >= (1LL << 32))
        return set_error(VFS_EINVAL, "vfs_unet_create: grid too large for the network's 32-bit channel indexing");
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
    long long wmax = (long long)OB_BLOCKS * (base + 1);
    for (auto *v : {&u->enc, &u->dec})
        for (auto &cv : *v) wmax = std::max({wmax, (long long)WG_SLICES * cv.nw, (long long)BG_BLOCKS * cv.cout});
    rc = rc ? rc : ualloc(u, (void **)&u->wpart, wmax * sizeof(float));
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
    return forward(u, b_dev, y_dev, false);
}

int vfs_unet_gradients(vfs_unet *u, const double *b_dev, const double *phi_dev, double *loss)
{
    if (!u || !b_dev || !phi_dev) return set_error(VFS_EINVAL, "vfs_unet_gradients: null argument");
    int rc = forward(u, b_dev, nullptr, true);
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
