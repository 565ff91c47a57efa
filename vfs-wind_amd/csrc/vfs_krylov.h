// vfs_krylov.h -- the Gram-Schmidt sweeps over a Krylov basis shared by the
// Poisson FGMRES and the implicit momentum GMRES.  Deterministic: per-block
// partials in a fixed order, reduced by k_finish_many.
#pragma once

#include <hip/hip_runtime.h>

#include "vfs_reduce.h"

namespace vfs {

// The Gram-Schmidt sweeps over the Krylov basis.  Each is instantiated for a
// maximum basis count NVM (4, 8, 16 or 32; the host picks the smallest that
// holds nv), so that the per-vector accumulators stay in registers without
// spills, and each thread keeps U grid-stride elements in flight.

// dots[v] = <w, V_v> for v < nv and dots[nv] = <w, w> (per-block partials)
template <int NVM, int U>
__global__ void __launch_bounds__(RED_THREADS) k_multidot(long long n, int nv, const double *__restrict__ V, long long ld,
                                                          const double *__restrict__ w, double *__restrict__ part)
{
    double acc[NVM];
#pragma unroll
    for (int v = 0; v < NVM; v++) acc[v] = 0;
    double ww = 0;
    const long long S = (long long)gridDim.x * blockDim.x;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; t + (U - 1) * S < n; t += U * S) {
        double wt[U];
#pragma unroll
        for (int u = 0; u < U; u++) wt[u] = w[t + u * S];
#pragma unroll
        for (int v = 0; v < NVM; v++)
            if (v < nv) {
                double vt[U];
#pragma unroll
                for (int u = 0; u < U; u++) vt[u] = V[v * ld + t + u * S];
#pragma unroll
                for (int u = 0; u < U; u++) acc[v] += vt[u] * wt[u];
            }
#pragma unroll
        for (int u = 0; u < U; u++) ww += wt[u] * wt[u];
    }
    for (; t < n; t += S) {
        const double wt = w[t];
#pragma unroll
        for (int v = 0; v < NVM; v++)
            if (v < nv) acc[v] += V[v * ld + t] * wt;
        ww += wt * wt;
    }
#pragma unroll
    for (int v = 0; v < NVM; v++) {
        if (v < nv) {
            const double r = block_reduce<OpSum>(acc[v]);
            if (threadIdx.x == 0) part[v * RED_BLOCKS + blockIdx.x] = r;
        }
    }
    const double r = block_reduce<OpSum>(ww);
    if (threadIdx.x == 0) part[nv * RED_BLOCKS + blockIdx.x] = r;
}

// w = w - sum_v h[v] V_v; with DOTS the per-block partials of <V_v, w_new>
// (v < nv), the dot products a second (re-orthogonalisation) pass needs, in
// the same sweep; always the partials of <w_new, w_new> (at slot nv with
// DOTS, slot 0 without)
template <int NVM, int U, bool DOTS>
__global__ void __launch_bounds__(RED_THREADS) k_multiaxpy(long long n, int nv, const double *__restrict__ h,
                                                           const double *__restrict__ V, long long ld,
                                                           double *__restrict__ w, double *__restrict__ part,
                                                           const double *__restrict__ gww, const double *__restrict__ gww0)
{
    // gate: run only when *gww < *gww0 / 4 (the re-orthogonalisation test)
    if (gww && !(gww[0] < 0.25 * gww0[0])) return;
    double hv[NVM], acc[NVM];
#pragma unroll
    for (int v = 0; v < NVM; v++) {
        hv[v] = v < nv ? h[v] : 0.0;
        acc[v] = 0;
    }
    double ss = 0;
    const long long S = (long long)gridDim.x * blockDim.x;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; t + (U - 1) * S < n; t += U * S) {
        double s[U], vt[NVM][U];
#pragma unroll
        for (int u = 0; u < U; u++) s[u] = w[t + u * S];
#pragma unroll
        for (int v = 0; v < NVM; v++)
            if (v < nv)
#pragma unroll
                for (int u = 0; u < U; u++) vt[v][u] = V[v * ld + t + u * S];
#pragma unroll
        for (int v = 0; v < NVM; v++)
            if (v < nv)
#pragma unroll
                for (int u = 0; u < U; u++) s[u] -= hv[v] * vt[v][u];
#pragma unroll
        for (int u = 0; u < U; u++) {
            w[t + u * S] = s[u];
            ss += s[u] * s[u];
        }
        if (DOTS)
#pragma unroll
            for (int v = 0; v < NVM; v++)
                if (v < nv)
#pragma unroll
                    for (int u = 0; u < U; u++) acc[v] += vt[v][u] * s[u];
    }
    for (; t < n; t += S) {
        double vt[NVM];
        double s = w[t];
#pragma unroll
        for (int v = 0; v < NVM; v++)
            if (v < nv) {
                vt[v] = V[v * ld + t];
                s -= hv[v] * vt[v];
            }
        w[t] = s;
        ss += s * s;
        if (DOTS)
#pragma unroll
            for (int v = 0; v < NVM; v++)
                if (v < nv) acc[v] += vt[v] * s;
    }
    if (DOTS) {
#pragma unroll
        for (int v = 0; v < NVM; v++) {
            if (v < nv) {
                const double r = block_reduce<OpSum>(acc[v]);
                if (threadIdx.x == 0) part[v * RED_BLOCKS + blockIdx.x] = r;
            }
        }
    }
    ss = block_reduce<OpSum>(ss);
    if (threadIdx.x == 0) part[(DOTS ? nv : 0) * RED_BLOCKS + blockIdx.x] = ss;
}

static __global__ void k_finish_many(int nv, const double *__restrict__ part, double *__restrict__ out)
{
    // one block per vector
    int v = blockIdx.x;
    if (v > nv) return;
    double s = 0;
    for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x) s += part[v * RED_BLOCKS + q];
    s = block_reduce<OpSum>(s);
    if (threadIdx.x == 0) out[v] = s;
}

}  // namespace vfs
