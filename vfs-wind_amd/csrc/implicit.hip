// implicit.hip -- Crank-Nicolson momentum step, -imp 4 (implicitsolver.c:4203-4446).
//
// Residual (FormFunction_SNES, momentum.c:2237-2336, time_coeff() == 1):
//   R(U) = -(U - U_o)/dt + 0.5 F(U) + 0.5 F(U_o) - dP (+ F_eul)
// with F = Formfunction_2 after the wall-flux zeroing and Contra2Cart of U.
// The reference hands R to PETSc's SNES (trust region, matrix-free finite
// difference Jacobian, GMRES without preconditioner, Eisenstat-Walker
// forcing).  Here the same solve runs on the device: an inexact Newton
// iteration with Eisenstat-Walker forcing and residual-decrease step
// halving, GMRES(m) with classical Gram-Schmidt and one re-orthogonalisation
// pass, and the matrix-free product J v = (R(U + h v) - R(U)) / h with
// PETSc's default differencing parameter.  Every vector (3 doubles per DA
// point) stays in HBM; the host sees only dot products.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"
#include "vfs_krylov.h"

#pragma clang fp contract(off)

namespace vfs {

constexpr int IM_TPB = 256;
constexpr int IM_MAXV = 33;  // GMRES basis vectors held by the dot kernel

// wall-flux zeroing of FormFunction_SNES (momentum.c:2258-2282)
__global__ void k_snes_bc(Grid g, double *__restrict__ ucont)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int i = t % g.mx;
    const long long r = t / g.mx;
    const int j = r % g.my, k = r / g.my;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const int *b = g.bc;
#define UC(kk, jj, ii, c) ucont[3 * (((long long)(kk) * my + (jj)) * mx + (ii)) + (c)]
    if (i == 0 && b[0] == 1) UC(k, j, i, 0) = 0;
    if (i == mx - 1 && b[1] == 1) UC(k, j, i - 1, 0) = 0;
    if (j == 0 && b[2] == 1) UC(k, j, i, 1) = 0;
    if (j == my - 1 && b[3] == 1) UC(k, j - 1, i, 1) = 0;
    if (k == 0 && b[4] == 1) UC(k, j, i, 2) = 0;
    if (k == mz - 1 && b[5] == 1) UC(k - 1, j, i, 2) = 0;
    if (j == my - 1 && b[3] == 2) UC(k, j - 1, i, 1) = 0;
    if (j == 0 && b[2] == 12) UC(k, j, i, 1) = 0;
    if (j == my - 1 && b[3] == 12) UC(k, j - 1, i, 1) = 0;
    if (b[0] == 10 && i == 0 && (j != 0 && j != my - 1 && k != 0 && k != mz - 1)) UC(k, j, i, 0) = 0;
    if (b[1] == 10 && i == mx - 1 && (j != 0 && j != my - 1 && k != 0 && k != mz - 1)) UC(k, j, i - 1, 0) = 0;
    if (b[2] == 10 && j == 0 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1)) UC(k, j, i, 1) = 0;
    if ((b[3] == 10 || b[3] == -10) && j == my - 1 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1)) UC(k, j - 1, i, 1) = 0;
#undef UC
}

// Rhs = 0; Rhs += (-1/dt) U; Rhs += (1/dt) U_o   (PETSc VecAXPY order)
__global__ void k_cn_base(long long n3, double rdt, const double *__restrict__ u, const double *__restrict__ uo,
                          double *__restrict__ rhs)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n3) return;
    double s = 0.0 + (-rdt) * u[t];
    rhs[t] = s + rdt * uo[t];
}

// R = Rhs + 0.5 RHS_o - dP (+ F_eul)
__global__ void k_cn_finish(long long n3, const double *__restrict__ rhs, const double *__restrict__ rhso,
                            const double *__restrict__ dp, const double *__restrict__ feul, double *__restrict__ r)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n3) return;
    double s = rhs[t] + 0.5 * rhso[t];
    s = s + (-1.0) * dp[t];
    if (feul) s = s + feul[t];
    r[t] = s;
}

// post-solve boundary fix of Implicit_MatrixFree (implicitsolver.c:4330-4392)
// for wall (1), slip (10) and outlet (4) faces
__global__ void k_imp_post(Grid g, const double *__restrict__ eta, const double *__restrict__ zet,
                           double *__restrict__ ucont, double *__restrict__ ucat)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int i = t % g.mx;
    const long long r = t / g.mx;
    const int j = r % g.my, k = r / g.my;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const int *b = g.bc;
    auto wall = [](int c) { return c == 1 || c == -1 || c == -2 || c == 10; };
    double *uc = ucont + 3 * t;
    if (wall(b[0]) && i == 0) uc[0] = 0;
    if (wall(b[1]) && i == mx - 2) uc[0] = 0;
    if (wall(b[2]) && j == 0) uc[1] = 0;
    if ((wall(b[3]) || b[3] == 2) && j == my - 2) uc[1] = 0;
    if (wall(b[4]) && k == 0) uc[2] = 0;
    if (wall(b[5]) && k == mz - 2) uc[2] = 0;
    if ((b[0] == -1 || b[0] == -2) && i == 1) uc[0] = 0;
    if ((b[1] == -1 || b[1] == -2) && i == mx - 3) uc[0] = 0;
    if ((b[2] == -1 || b[2] == -2) && j == 1) uc[1] = 0;
    if ((b[3] == -1 || b[3] == -2) && j == my - 3) uc[1] = 0;
    if ((b[4] == -1 || b[4] == -2) && k == 1) uc[2] = 0;
    if ((b[5] == -1 || b[5] == -2) && k == mz - 3) uc[2] = 0;
    if (b[3] == 4 && j == my - 2) {
        const long long q1 = t + mx;
        for (int c = 0; c < 3; c++) ucat[3 * q1 + c] = ucat[3 * t + c];
        const double *e = eta + 3 * t;
        uc[1] = 0.5 * (ucat[3 * t] + ucat[3 * q1]) * e[0] + 0.5 * (ucat[3 * t + 1] + ucat[3 * q1 + 1]) * e[1] +
                0.5 * (ucat[3 * t + 2] + ucat[3 * q1 + 2]) * e[2];
    }
    if (b[5] == 4 && k == mz - 2) {
        const long long q1 = t + g.plane;
        for (int c = 0; c < 3; c++) ucat[3 * q1 + c] = ucat[3 * t + c];
        const double *z = zet + 3 * t;
        uc[2] = 0.5 * (ucat[3 * t] + ucat[3 * q1]) * z[0] + 0.5 * (ucat[3 * t + 1] + ucat[3 * q1 + 1]) * z[1] +
                0.5 * (ucat[3 * t + 2] + ucat[3 * q1 + 2]) * z[2];
        if (uc[2] < 0) {
            uc[2] = 0;
            for (int c = 0; c < 3; c++) ucat[3 * t + c] = ucat[3 * q1 + c] = 0;
        }
    }
}

// ---- device vector algebra (deterministic reductions; the Gram-Schmidt
// sweeps are vfs_krylov.h's) --------------------------------------------------
// y = a x + b y, coefficients from device memory (a = s[ia], b = s[ib] or 0/1)
__global__ void k_im_axpby(long long n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = a * x[t] + b * y[t];
}

// z = x + a y
__global__ void k_im_waxpy(long long n, const double *__restrict__ x, double a, const double *__restrict__ y,
                           double *__restrict__ z)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    z[t] = x[t] + a * y[t];
}

// x = sum_v y[v] V_v (first) or x + sum_v y[v] V_v, the terms added in v
// order with the fused multiply-add of a chain of k_im_axpby updates
__global__ void k_im_update(long long n, int nv, const double *__restrict__ y, const double *__restrict__ V, long long ld,
                            double *__restrict__ x, int first)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = first ? 0.0 : x[t];
    for (int v = 0; v < nv; v++) s = y[v] * V[v * ld + t] + s;
    x[t] = s;
}

// z = z2 = x + a y (the line search's trial state and the residual's input)
__global__ void k_im_waxpy2(long long n, const double *__restrict__ x, double a, const double *__restrict__ y,
                            double *__restrict__ z, double *__restrict__ z2)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = x[t] + a * y[t];
    z[t] = v;
    z2[t] = v;
}

// sum of |x| (for the differencing parameter)
__global__ void __launch_bounds__(RED_THREADS) k_im_asum_dot(long long n, const double *__restrict__ u,
                                                             const double *__restrict__ v, double *__restrict__ part)
{
    double a = 0, d = 0, vv = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double x = v[t];
        a += fabs(x);
        d += u[t] * x;
        vv += x * x;
    }
    a = block_reduce<OpSum>(a);
    d = block_reduce<OpSum>(d);
    vv = block_reduce<OpSum>(vv);
    if (threadIdx.x == 0) { part[blockIdx.x] = a; part[RED_BLOCKS + blockIdx.x] = d; part[2 * RED_BLOCKS + blockIdx.x] = vv; }
}

}  // namespace vfs

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace vfs {

struct ImpState {
    long long n3 = 0;
    int m = 0;
    double *U = nullptr, *R = nullptr, *W = nullptr, *FW = nullptr, *dx = nullptr, *V = nullptr, *h = nullptr;
};

// Carve the solver's vectors out of the context's persistent workspace
// (5 + m + 1 vectors of n3 doubles and IM_MAXV scalars), allocating or
// growing it first when needed.  Every vector is written before it is read.
static int im_workspace(vfs_ctx *c, ImpState &s)
{
    const size_t vb = (size_t)s.n3 * sizeof(double);
    const size_t need = (size_t)(6 + s.m) * vb + IM_MAXV * sizeof(double);
    if (c->imp_ws_bytes < need) {
        VFS_HIP(hipStreamSynchronize(c->stream));
        if (c->imp_ws) hipFree(c->imp_ws);
        c->imp_ws = nullptr;
        c->imp_ws_bytes = 0;
        hipError_t e = hipMalloc((void **)&c->imp_ws, need);
        if (e != hipSuccess) { c->imp_ws = nullptr; return check_hip(e, "hipMalloc (implicit)"); }
        e = hipMemsetAsync(c->imp_ws, 0, need, c->stream);
        if (e != hipSuccess) return check_hip(e, "hipMemsetAsync (implicit)");
        c->imp_ws_bytes = need;
    }
    double *p = c->imp_ws;
    s.U = p; p += s.n3;
    s.R = p; p += s.n3;
    s.W = p; p += s.n3;
    s.FW = p; p += s.n3;
    s.dx = p; p += s.n3;
    s.V = p; p += (size_t)(s.m + 1) * s.n3;
    s.h = p;
    return VFS_OK;
}

// R(u) into r (FormFunction_SNES); leaves UCONT/UCAT at the BC-zeroed u.
// With rbase, r = da rbase + db R(u) (one k_im_axpby's arithmetic), formed in
// the cell pass on the fused path.
int snes_function(vfs_ctx *c, const double *u, double *r, int use_feul, const double *rbase = nullptr, double da = 0,
                  double db = 0)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    const long long n3 = 3 * g.n;
    if (u != c->f[VFS_UCONT])
        VFS_HIP(hipMemcpyAsync(c->f[VFS_UCONT], u, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_snes_bc, dim3(blocks_for(g.n, IM_TPB)), dim3(IM_TPB), 0, st, g, c->f[VFS_UCONT]);
    ucat_stale(c);
    VFS_LAUNCH_CHECK("snes bc");
    int rc = c2c(c);
    if (rc) return rc;
    // on the fused path the residual is formed in the cell pass (the same
    // operations in the same order); VFS_FUSED_CN=0 keeps the three passes
    const char *fe = std::getenv("VFS_FUSED_CN");
    if (formfunction2_fused(c) && !(fe && fe[0] == '0')) {
        const CnResidual cn{c->f[VFS_UCONT_O], c->f[VFS_RHS_O], use_feul ? c->f[VFS_FEUL] : nullptr, 1. / c->p.dt, r,
                            rbase, da, db};
        return formfunction2(c, 0.5, nullptr, 0.0, nullptr, &cn);
    }
    hipLaunchKernelGGL(k_cn_base, dim3(blocks_for(n3, IM_TPB)), dim3(IM_TPB), 0, st, n3, 1. / c->p.dt, c->f[VFS_UCONT],
                       c->f[VFS_UCONT_O], c->f[VFS_RHS]);
    rc = formfunction2(c, 0.5, nullptr, 0.0, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cn_finish, dim3(blocks_for(n3, IM_TPB)), dim3(IM_TPB), 0, st, n3, c->f[VFS_RHS], c->f[VFS_RHS_O],
                       c->f[VFS_DP], use_feul ? c->f[VFS_FEUL] : (const double *)nullptr, r);
    if (rbase) hipLaunchKernelGGL(k_im_axpby, dim3(blocks_for(n3, IM_TPB)), dim3(IM_TPB), 0, st, n3, da, rbase, db, r);
    VFS_LAUNCH_CHECK("snes residual");
    return VFS_OK;
}

// the nv + 1 sums of the partials in c->red, on the host in out
static int im_finish(vfs_ctx *c, int nv, double *out)
{
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_finish_many, dim3(nv + 1), dim3(RED_THREADS), 0, st, nv, c->red, c->red + IM_MAXV * RED_BLOCKS);
    VFS_LAUNCH_CHECK("implicit dots");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + IM_MAXV * RED_BLOCKS, (nv + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    std::memcpy(out, c->red_host, (nv + 1) * sizeof(double));
    return VFS_OK;
}

// <w, V_v> for v < nv and <w, w> on the host
static int im_dots(vfs_ctx *c, long long n, int nv, const double *V, const double *w, double *out)
{
    hipStream_t st = c->stream;
    const dim3 gb(RED_BLOCKS), tb(RED_THREADS);
    if (nv <= 4) hipLaunchKernelGGL((k_multidot<4, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 8) hipLaunchKernelGGL((k_multidot<8, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 16) hipLaunchKernelGGL((k_multidot<16, 2>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else hipLaunchKernelGGL((k_multidot<IM_MAXV - 1, 1>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    return im_finish(c, nv, out);
}

// w -= sum_v h_v V_v (h on the device in s.h); with DOTS the partials of
// <V_v, w_new> (v < nv) and <w_new, w_new>, else of <w_new, w_new> alone;
// the finished sums on the host in out
template <bool DOTS>
static int im_multiaxpy(vfs_ctx *c, long long n, int nv, const double *h, const double *V, double *w, double *out)
{
    hipStream_t st = c->stream;
    const dim3 gb(RED_BLOCKS), tb(RED_THREADS);
    const double *g0 = nullptr;
    if (nv <= 4) hipLaunchKernelGGL((k_multiaxpy<4, 4, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, g0, g0);
    else if (nv <= 8) hipLaunchKernelGGL((k_multiaxpy<8, 4, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, g0, g0);
    else if (nv <= 16) hipLaunchKernelGGL((k_multiaxpy<16, 2, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, g0, g0);
    else hipLaunchKernelGGL((k_multiaxpy<IM_MAXV - 1, 1, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, g0, g0);
    return im_finish(c, DOTS ? nv : 0, out);
}

static int im_norm(vfs_ctx *c, long long n, const double *w, double *out)
{
    double d[1];
    int rc = im_dots(c, n, 0, nullptr, w, d);
    *out = std::sqrt(d[0]);
    return rc;
}

// J v = (R(U + h v) - R(U)) / h, PETSc MatMFFD "ds" parameter:
// h = e * (u.v) / |v|^2, or e * umin * sign(u.v) * |v|_1 / |v|^2 when |u.v| < umin |v|_1
static int mffd(vfs_ctx *c, ImpState &s, const double *v, double *jv, int use_feul, int *nfunc)
{
    hipStream_t st = c->stream;
    const long long n = s.n3;
    hipLaunchKernelGGL(k_im_asum_dot, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, s.U, v, c->red);
    for (int q = 0; q < 3; q++)
        hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red + q * RED_BLOCKS, RED_BLOCKS,
                           c->red + 3 * RED_BLOCKS + q);
    VFS_LAUNCH_CHECK("mffd norms");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + 3 * RED_BLOCKS, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    const double asum = c->red_host[0], dot = c->red_host[1], vv = c->red_host[2];
    if (vv == 0) {
        VFS_HIP(hipMemsetAsync(jv, 0, n * sizeof(double), st));
        return VFS_OK;
    }
    const double err = 1.490116119384766e-08, umin = 1.e-6;  // sqrt(DBL_EPSILON), PETSc defaults
    double hh = dot;
    if (std::fabs(hh) < umin * asum) hh = (hh >= 0 ? 1.0 : -1.0) * umin * asum;
    hh = err * hh / vv;
    // U + h v goes straight into UCONT, where the residual evaluation reads it
    hipLaunchKernelGGL(k_im_waxpy, dim3(blocks_for(n, IM_TPB)), dim3(IM_TPB), 0, st, n, s.U, hh, v, c->f[VFS_UCONT]);
    // jv = (F(U + h v) - F(U)) / h, the quotient formed with the residual
    int rc = snes_function(c, c->f[VFS_UCONT], jv, use_feul, s.R, -1.0 / hh, 1.0 / hh);
    if (rc) return rc;
    (*nfunc)++;
    VFS_LAUNCH_CHECK("mffd");
    return VFS_OK;
}

// GMRES(m) for J dx = -R, dx from zero; returns iterations, final estimate
static int gmres(vfs_ctx *c, ImpState &s, double rtol, int maxit, int use_feul, int *its, int *nfunc)
{
    hipStream_t st = c->stream;
    const long long n = s.n3;
    const int m = s.m;
    const unsigned nb = blocks_for(n, IM_TPB);
    double d[IM_MAXV + 1];
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    // r0 = -R (dx = 0); dx is first written by the first restart's update
    int rc = im_dots(c, n, 0, nullptr, s.R, d);
    if (rc) return rc;
    const double bnorm = std::sqrt(d[0]);
    *its = 0;
    if (bnorm == 0) {
        VFS_HIP(hipMemsetAsync(s.dx, 0, n * sizeof(double), st));
        return VFS_OK;
    }
    double beta = bnorm;
    bool first = true, restart0 = true;
    while (*its < maxit) {
        double *V0 = s.V;
        if (first) {
            hipLaunchKernelGGL(k_im_axpby, dim3(nb), dim3(IM_TPB), 0, st, n, -1.0 / beta, s.R, 0.0, V0);
        } else {
            // r = -R - J dx
            rc = mffd(c, s, s.dx, s.FW, use_feul, nfunc);
            if (rc) return rc;
            hipLaunchKernelGGL(k_im_axpby, dim3(nb), dim3(IM_TPB), 0, st, n, -1.0, s.R, -1.0, s.FW);
            rc = im_dots(c, n, 0, nullptr, s.FW, d);
            if (rc) return rc;
            beta = std::sqrt(d[0]);
            if (beta <= rtol * bnorm) break;
            hipLaunchKernelGGL(k_im_axpby, dim3(nb), dim3(IM_TPB), 0, st, n, 1.0 / beta, s.FW, 0.0, V0);
        }
        first = false;
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int jj;
        for (jj = 0; jj < m && *its < maxit; jj++, (*its)++) {
            double *vj = s.V + (size_t)jj * n, *w = s.V + (size_t)(jj + 1) * n;
            rc = mffd(c, s, vj, w, use_feul, nfunc);
            if (rc) return rc;
            // classical Gram-Schmidt; the first sweep w -= V h also forms the
            // dot products a second (re-orthogonalisation) pass needs
            double h[IM_MAXV];
            rc = im_dots(c, n, jj + 1, s.V, w, d);
            if (rc) return rc;
            for (int v = 0; v <= jj; v++) h[v] = d[v];
            const double ww0 = d[jj + 1];
            std::memcpy(c->red_host, h, (jj + 1) * sizeof(double));
            VFS_HIP(hipMemcpyAsync(s.h, c->red_host, (jj + 1) * sizeof(double), hipMemcpyHostToDevice, st));
            rc = im_multiaxpy<true>(c, n, jj + 1, s.h, s.V, w, d);
            if (rc) return rc;
            double ww = d[jj + 1];
            if (ww < 0.25 * ww0) {  // second classical Gram-Schmidt pass
                for (int v = 0; v <= jj; v++) h[v] += d[v];
                std::memcpy(c->red_host, d, (jj + 1) * sizeof(double));
                VFS_HIP(hipMemcpyAsync(s.h, c->red_host, (jj + 1) * sizeof(double), hipMemcpyHostToDevice, st));
                double d2[IM_MAXV + 1];
                rc = im_multiaxpy<false>(c, n, jj + 1, s.h, s.V, w, d2);
                if (rc) return rc;
                ww = d2[0];
            }
            const double hn = std::sqrt(ww > 0 ? ww : 0);
            for (int v = 0; v <= jj; v++) H[(size_t)v * m + jj] = h[v];
            H[(size_t)(jj + 1) * m + jj] = hn;
            if (hn > 0)
                hipLaunchKernelGGL(k_im_axpby, dim3(nb), dim3(IM_TPB), 0, st, n, 1.0 / hn, w, 0.0, w);
            for (int i = 0; i < jj; i++) {
                const double tmp = cs[i] * H[(size_t)i * m + jj] + sn[i] * H[(size_t)(i + 1) * m + jj];
                H[(size_t)(i + 1) * m + jj] = -sn[i] * H[(size_t)i * m + jj] + cs[i] * H[(size_t)(i + 1) * m + jj];
                H[(size_t)i * m + jj] = tmp;
            }
            const double a = H[(size_t)jj * m + jj], bb = H[(size_t)(jj + 1) * m + jj];
            const double den = std::sqrt(a * a + bb * bb);
            cs[jj] = den > 0 ? a / den : 1.0;
            sn[jj] = den > 0 ? bb / den : 0.0;
            H[(size_t)jj * m + jj] = den;
            H[(size_t)(jj + 1) * m + jj] = 0;
            gv[jj + 1] = -sn[jj] * gv[jj];
            gv[jj] = cs[jj] * gv[jj];
            if (std::fabs(gv[jj + 1]) <= rtol * bnorm || hn == 0) { jj++; (*its)++; break; }
        }
        for (int i = jj - 1; i >= 0; i--) {
            double sacc = gv[i];
            for (int q = i + 1; q < jj; q++) sacc -= H[(size_t)i * m + q] * y[q];
            y[i] = H[(size_t)i * m + i] != 0 ? sacc / H[(size_t)i * m + i] : 0.0;
        }
        // dx += sum_q y_q V_q in one sweep (dx = 0 before the first restart)
        std::memcpy(c->red_host, y.data(), jj * sizeof(double));
        VFS_HIP(hipMemcpyAsync(s.h, c->red_host, jj * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_im_update, dim3(nb), dim3(IM_TPB), 0, st, n, jj, s.h, s.V, n, s.dx, restart0 ? 1 : 0);
        restart0 = false;
        VFS_LAUNCH_CHECK("gmres update");
        if (std::fabs(gv[jj]) <= rtol * bnorm) break;
    }
    if (restart0) VFS_HIP(hipMemsetAsync(s.dx, 0, n * sizeof(double), st));  // no update was made
    return VFS_OK;
}

}  // namespace vfs

using namespace vfs;

extern "C" {

int vfs_formfunction_snes(vfs_ctx *c, const double *u_dev, double *r_dev, int use_feul)
{
    if (!c || !u_dev || !r_dev) return set_error(VFS_EINVAL, "vfs_formfunction_snes: null argument");
    return snes_function(c, u_dev, r_dev, use_feul);
}

int vfs_store_old_state(vfs_ctx *c)
{
    if (!c) return set_error(VFS_EINVAL, "null context");
    const size_t n3 = 3 * (size_t)c->g.n;
    // Ucont_o = Ucont; RHS_o = Formfunction_2(Ucont) (solvers.c:627-628)
    VFS_HIP(hipMemcpyAsync(c->f[VFS_UCONT_O], c->f[VFS_UCONT], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    VFS_HIP(hipMemsetAsync(c->f[VFS_RHS], 0, n3 * sizeof(double), c->stream));
    int rc = formfunction2(c, 1.0, nullptr, 0.0, nullptr);
    if (rc) return rc;
    VFS_HIP(hipMemcpyAsync(c->f[VFS_RHS_O], c->f[VFS_RHS], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return VFS_OK;
}

int vfs_implicit_matrixfree(vfs_ctx *c, const vfs_newton_opts *o, vfs_newton_stats *stats)
{
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    if (!c || !o) return set_error(VFS_EINVAL, "vfs_implicit_matrixfree: null argument");
    const int m = o->restart > 0 ? (o->restart < IM_MAXV - 1 ? o->restart : IM_MAXV - 2) : 30;
    ImpState s;
    s.n3 = 3 * c->g.n;
    s.m = m;
    const size_t vb = (size_t)s.n3 * sizeof(double);
    int rc = im_workspace(c, s);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const long long n = s.n3;
    const unsigned nb = blocks_for(n, IM_TPB);
    auto fail = [&](int code) { return code; };
    // InflowFlux, then FormBCS with outflow_scale = 0 (implicitsolver.c:4293-4297)
    rc = inflow_flux(c, nullptr);
    if (!rc) rc = form_bcs(c, 0);
    if (rc) return fail(rc);
    VFS_HIP(hipMemcpyAsync(s.U, c->f[VFS_UCONT], vb, hipMemcpyDeviceToDevice, st));
    int nfunc = 0, newton = 0, lin_total = 0;
    rc = snes_function(c, s.U, s.R, o->use_feul);
    if (rc) return fail(rc);
    nfunc++;
    double fnorm;
    rc = im_norm(c, n, s.R, &fnorm);
    if (rc) return fail(rc);
    const double fnorm0 = fnorm;
    const double rtol = o->rtol > 0 ? o->rtol : 1e-4;
    const int maxit = o->max_newton > 0 ? o->max_newton : 50;
    const int maxlin = o->max_linear > 0 ? o->max_linear : 1000;
    // Eisenstat-Walker, version 3 parameters (implicitsolver.c:4251-4252, PETSc defaults)
    const double ew_rtol0 = 0.3, ew_max = 0.9, ew_gamma = 1.0, ew_alpha = 0.5 * (1.0 + std::sqrt(5.0)), ew_thresh = 0.1;
    double eta = ew_rtol0, fnorm_last = fnorm;
    bool converged = fnorm <= rtol * fnorm0 || fnorm == 0;
    while (!converged && newton < maxit) {
        if (newton > 0) {
            double e = ew_gamma * std::pow(fnorm / fnorm_last, ew_alpha);
            const double sg = ew_gamma * std::pow(eta, ew_alpha);
            if (sg > ew_thresh) e = std::fmax(e, sg);
            eta = std::fmin(e, ew_max);
        }
        int its = 0;
        rc = gmres(c, s, eta, maxlin, o->use_feul, &its, &nfunc);
        if (rc) return fail(rc);
        lin_total += its;
        // step with halving while the residual does not decrease
        double lam = 1.0, fnew = 0;
        bool accepted = false;
        for (int half = 0; half < 8; half++) {
            hipLaunchKernelGGL(k_im_waxpy2, dim3(nb), dim3(IM_TPB), 0, st, n, s.U, lam, s.dx, s.W, c->f[VFS_UCONT]);
            rc = snes_function(c, c->f[VFS_UCONT], s.FW, o->use_feul);
            if (rc) return fail(rc);
            nfunc++;
            rc = im_norm(c, n, s.FW, &fnew);
            if (rc) return fail(rc);
            if (fnew < fnorm) { accepted = true; break; }
            lam *= 0.5;
        }
        newton++;
        if (!accepted) break;
        std::swap(s.U, s.W);
        std::swap(s.R, s.FW);
        fnorm_last = fnorm;
        fnorm = fnew;
        converged = fnorm <= rtol * fnorm0;
    }
    // U -> Ucont with Ucat as the last residual evaluation at U left it, the
    // post-solve boundary fix and FormBCS(outflow_scale = 1)
    VFS_HIP(hipMemcpyAsync(c->f[VFS_UCONT], s.U, vb, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_snes_bc, dim3(blocks_for(c->g.n, IM_TPB)), dim3(IM_TPB), 0, st, c->g, c->f[VFS_UCONT]);
    ucat_stale(c);
    rc = c2c(c);
    if (rc) return fail(rc);
    hipLaunchKernelGGL(k_imp_post, dim3(blocks_for(c->g.n, IM_TPB)), dim3(IM_TPB), 0, st, c->g, c->f[VFS_ETA], c->f[VFS_ZET],
                       c->f[VFS_UCONT], c->f[VFS_UCAT]);
    ucat_stale(c);
    VFS_LAUNCH_CHECK("implicit post");
    rc = form_bcs(c, 1);
    if (rc) return fail(rc);
    VFS_HIP(hipStreamSynchronize(st));
    if (stats) {
        stats->newton_iterations = newton;
        stats->linear_iterations = lin_total;
        stats->function_evals = nfunc;
        stats->fnorm0 = fnorm0;
        stats->fnorm = fnorm;
        stats->ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    if (!converged) return set_error(VFS_ENOCONV, "implicit momentum: Newton did not reach the tolerance");
    return VFS_OK;
}

}  // extern "C"
