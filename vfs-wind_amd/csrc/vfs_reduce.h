// vfs_reduce.h -- deterministic FP64 reductions for CDNA4 (64-wide waves).
//
// Each block reduces a grid-stride range with a fixed tree (wave shuffles,
// then LDS across the block's waves); the per-block partials are reduced by a
// single block in a second launch.  The order never depends on timing, so a
// result is bitwise reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>

namespace vfs {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;  // partials per reduction (fits ctx->red)

struct OpSum {
    __device__ static double init() { return 0.0; }
    __device__ static double op(double a, double b) { return a + b; }
};
struct OpMax {
    __device__ static double init() { return -1.0e308; }
    __device__ static double op(double a, double b) { return a > b ? a : b; }
};
struct OpMin {
    __device__ static double init() { return 1.0e308; }
    __device__ static double op(double a, double b) { return a < b ? a : b; }
};

template <class Op>
__device__ inline double wave_reduce(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = Op::op(v, __shfl_down(v, off, 64));
    return v;
}

// Reduces v over the block; the result is valid in thread 0.
template <class Op>
__device__ inline double block_reduce(double v)
{
    __shared__ double part[RED_THREADS / 64];
    v = wave_reduce<Op>(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double r = Op::init();
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) r = Op::op(r, part[q]);
    __syncthreads();
    return r;
}

// Second stage: reduce `n` partials (n <= RED_BLOCKS * k) into out[0].
template <class Op>
__global__ void __launch_bounds__(RED_THREADS) finish_reduce(const double *__restrict__ part, int n,
                                                             double *__restrict__ out)
{
    double v = Op::init();
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = Op::op(v, part[q]);
    v = block_reduce<Op>(v);
    if (threadIdx.x == 0) out[0] = v;
}

// The same, adding the result to out[0] (partial results of several passes
// summed in launch order).
template <class Op>
__global__ void __launch_bounds__(RED_THREADS) finish_reduce_add(const double *__restrict__ part, int n,
                                                                 double *__restrict__ out)
{
    double v = Op::init();
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = Op::op(v, part[q]);
    v = block_reduce<Op>(v);
    if (threadIdx.x == 0) out[0] = Op::op(out[0], v);
}

}  // namespace vfs
