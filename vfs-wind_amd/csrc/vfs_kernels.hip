// vfs_kernels.hip -- metrics, Contra2Cart, FormBCS, LES eddy viscosity,
// Pressure_Gradient and the scalar diagnostics, as HIP kernels for gfx950.
//
// Every kernel is one pass over a DMDA array (i fastest, one thread per
// point, consecutive lanes on consecutive i) so loads and stores coalesce;
// the boundary work launches only over the boundary planes it touches.
#include <hip/hip_runtime.h>

#include <cmath>

#include "vfs_internal.h"
#include "vfs_reduce.h"

// FP64 expressions keep the reference's operation order: no contraction to FMA
#pragma clang fp contract(off)

namespace vfs {

constexpr int TPB = 256;

// ---------------------------------------------------------------------------
// periodic ghost fix over the two boundary planes of one periodic direction
// ---------------------------------------------------------------------------

// dir: 0 = i planes, 1 = j planes, 2 = k planes.  mode 0 copies all nc
// components; mode 1 (Projection, poisson.c:3001-3040) copies component d of
// a vector only when direction d is a periodic ghost index of the cell.
// One launch over the ghost planes of every periodic direction (t runs over
// the planes of the first periodic direction, then the next ...).  A ghost
// cell's source wraps all periodic directions at once (periodic_src), so it
// is never a periodic ghost itself and the planes do not depend on each
// other; a cell on two planes is written twice with the same value.
__global__ void k_periodic_planes(Grid g, double *__restrict__ f, int nc, int mode)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    const long long n0 = g.pi ? 2LL * g.my * g.mz : 0, n1 = g.pj ? 2LL * g.mx * g.mz : 0;
    const long long n2 = g.pk ? 2LL * g.mx * g.my : 0;
    if (t < n0) {
        int side = t & 1; long long r = t >> 1;
        j = r % g.my; k = r / g.my; i = side ? g.mx - 1 : 0;
    } else if ((t -= n0) < n1) {
        i = t % g.mx; long long r = t / g.mx;
        int side = r & 1; k = r >> 1; j = side ? g.my - 1 : 0;
    } else if ((t -= n1) < n2) {
        i = t % g.mx; long long r = t / g.mx;
        j = r % g.my; int side = (int)(r / g.my); k = side ? g.mz - 1 : 0;
    } else {
        return;
    }
    int a, b, c, fi, fj, fk;
    if (!periodic_src(g, k, j, i, c, b, a, fi, fj, fk)) return;
    long long d = gidx(g, k, j, i), s = gidx(g, c, b, a);
    if (mode == 0) {
        for (int q = 0; q < nc; q++) f[nc * d + q] = f[nc * s + q];
    } else {
        if (fi) f[3 * d + 0] = f[3 * s + 0];
        if (fj) f[3 * d + 1] = f[3 * s + 1];
        if (fk) f[3 * d + 2] = f[3 * s + 2];
    }
}

int periodic_fix_grid(hipStream_t st, const Grid &g, double *f, int nc, int mode)
{
    const long long n = (g.pi ? 2LL * g.my * g.mz : 0) + (g.pj ? 2LL * g.mx * g.mz : 0) + (g.pk ? 2LL * g.mx * g.my : 0);
    if (n == 0) return VFS_OK;
    hipLaunchKernelGGL(k_periodic_planes, dim3(blocks_for(n, TPB)), dim3(TPB), 0, st, g, f, nc, mode);
    VFS_LAUNCH_CHECK("periodic fix");
    return VFS_OK;
}

int periodic_fix(vfs_ctx *c, double *f, int nc, int mode)
{
    return periodic_fix_grid(c->stream, c->g, f, nc, mode);
}

// ---------------------------------------------------------------------------
// a1: FormMetrics (metrics.c:12-1355)
// ---------------------------------------------------------------------------

struct Decode {
    int i, j, k;
};
__device__ inline bool decode(const Grid &g, long long t, Decode &d)
{
    if (t >= g.n) return false;
    const int t32 = (int)t;  // g.n < 2^31 (checked when the context is created)
    d.i = t32 % g.mx;
    const int r = t32 / g.mx;
    d.j = r % g.my;
    d.k = r / g.my;
    return true;
}

#define CRD(c, kk, jj, ii) coor[3 * gidx(g, (kk), (jj), (ii)) + (c)]

// cell centres (metrics.c:79-100) and, at interior cells, the Jacobian with
// the cell metrics csi/eta/zet.  The reference compiles FormMetrics with
// NEWMETRIC defined (metrics.c:10): its node-plane formulas
// (metrics.c:115-191) are dead code, and the cell metrics are the cross
// products of the cell-averaged edge vectors (metrics.c:193-257).
__global__ void k_metrics_cell(Grid g, const double *__restrict__ coor, double *__restrict__ csi,
                               double *__restrict__ eta, double *__restrict__ zet, double *__restrict__ aj,
                               double *__restrict__ cent)
{
    Decode d;
    if (!decode(g, blockIdx.x * (long long)blockDim.x + threadIdx.x, d)) return;
    const int i = d.i, j = d.j, k = d.k;
    const long long q = gidx(g, k, j, i);
    if (!(i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1)) return;
    for (int c = 0; c < 3; c++) {
        double s = CRD(c, k, j, i) + CRD(c, k, j - 1, i) + CRD(c, k - 1, j, i) + CRD(c, k - 1, j - 1, i) +
                   CRD(c, k, j, i - 1) + CRD(c, k, j - 1, i - 1) + CRD(c, k - 1, j, i - 1) +
                   CRD(c, k - 1, j - 1, i - 1);
        cent[3 * q + c] = 0.125 * s;
    }
    double dd[3][3];
    for (int c = 0; c < 3; c++) {
        dd[c][0] = 0.25 * (CRD(c, k, j, i) + CRD(c, k, j - 1, i) + CRD(c, k - 1, j, i) + CRD(c, k - 1, j - 1, i) -
                           CRD(c, k, j, i - 1) - CRD(c, k, j - 1, i - 1) - CRD(c, k - 1, j, i - 1) -
                           CRD(c, k - 1, j - 1, i - 1));
        dd[c][1] = 0.25 * (CRD(c, k, j, i) + CRD(c, k, j, i - 1) + CRD(c, k - 1, j, i) + CRD(c, k - 1, j, i - 1) -
                           CRD(c, k, j - 1, i) - CRD(c, k, j - 1, i - 1) - CRD(c, k - 1, j - 1, i) -
                           CRD(c, k - 1, j - 1, i - 1));
        dd[c][2] = 0.25 * (CRD(c, k, j, i) + CRD(c, k, j - 1, i) + CRD(c, k, j, i - 1) + CRD(c, k, j - 1, i - 1) -
                           CRD(c, k - 1, j, i) - CRD(c, k - 1, j - 1, i) - CRD(c, k - 1, j, i - 1) -
                           CRD(c, k - 1, j - 1, i - 1));
    }
    const double dxdc = dd[0][0], dydc = dd[1][0], dzdc = dd[2][0];
    const double dxde = dd[0][1], dyde = dd[1][1], dzde = dd[2][1];
    const double dxdz = dd[0][2], dydz = dd[1][2], dzdz = dd[2][2];
    double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                 dzdc * (dxde * dydz - dyde * dxdz);
    aj[q] = 1. / det;
    csi[3 * q + 0] = dyde * dzdz - dzde * dydz;
    csi[3 * q + 1] = -dxde * dzdz + dzde * dxdz;
    csi[3 * q + 2] = dxde * dydz - dyde * dxdz;
    eta[3 * q + 0] = dydz * dzdc - dzdz * dydc;
    eta[3 * q + 1] = -dxdz * dzdc + dzdz * dxdc;
    eta[3 * q + 2] = dxdz * dydc - dydz * dxdc;
    zet[3 * q + 0] = dydc * dzde - dzdc * dyde;
    zet[3 * q + 1] = -dxdc * dzde + dzdc * dxde;
    zet[3 * q + 2] = dxdc * dyde - dydc * dxde;
}

// ghost mirroring (metrics.c:261-345): `which` selects one of the six
// sequential plane passes (i-, i+, j-, j+, k-, k+), each copying all four
// cell arrays from the neighbour plane; later passes overwrite edges.
__global__ void k_metrics_mirror(Grid g, int which, double *__restrict__ csi, double *__restrict__ eta,
                                 double *__restrict__ zet, double *__restrict__ aj)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int axis = which >> 1, hi = which & 1;
    int i, j, k;
    if (axis == 0) {
        if (t >= (long long)g.my * g.mz) return;
        j = t % g.my; k = t / g.my; i = hi ? g.mx - 1 : 0;
    } else if (axis == 1) {
        if (t >= (long long)g.mx * g.mz) return;
        i = t % g.mx; k = t / g.mx; j = hi ? g.my - 1 : 0;
    } else {
        if (t >= (long long)g.mx * g.my) return;
        i = t % g.mx; j = t / g.mx; k = hi ? g.mz - 1 : 0;
    }
    int si = i, sj = j, sk = k;
    if (axis == 0) si = hi ? g.mx - 2 : 1;
    if (axis == 1) sj = hi ? g.my - 2 : 1;
    if (axis == 2) sk = hi ? g.mz - 2 : 1;
    long long d = gidx(g, k, j, i), s = gidx(g, sk, sj, si);
    for (int c = 0; c < 3; c++) {
        csi[3 * d + c] = csi[3 * s + c];
        eta[3 * d + c] = eta[3 * s + c];
        zet[3 * d + c] = zet[3 * s + c];
    }
    aj[d] = aj[s];
}

// face metrics (metrics.c:588-593, 696-701, 804-809): the mean of the two
// cell metrics a face separates (AxByC, momentum.c:40-44) and the harmonic
// mean of their aj.  I faces cover i = 0..mx-2 over the interior j, k; J and
// K faces likewise in their own direction (metrics.c:497, 626, 731).
__global__ void k_face_metrics(Grid g, int axis, const double *__restrict__ csi, const double *__restrict__ eta,
                               const double *__restrict__ zet, const double *__restrict__ aj,
                               double *__restrict__ fcsi, double *__restrict__ feta, double *__restrict__ fzet,
                               double *__restrict__ faj)
{
    Decode d;
    if (!decode(g, blockIdx.x * (long long)blockDim.x + threadIdx.x, d)) return;
    const int i = d.i, j = d.j, k = d.k;
    const int i0 = axis == 0 ? 0 : 1, j0 = axis == 1 ? 0 : 1, k0 = axis == 2 ? 0 : 1;
    if (!(i >= i0 && i < g.mx - 1 && j >= j0 && j < g.my - 1 && k >= k0 && k < g.mz - 1)) return;
    const long long q = gidx(g, k, j, i);
    const long long r = q + (axis == 0 ? 1LL : axis == 1 ? (long long)g.mx : (long long)g.mx * g.my);
    for (int c = 0; c < 3; c++) {
        fcsi[3 * q + c] = 0.5 * csi[3 * q + c] + 0.5 * csi[3 * r + c];
        feta[3 * q + c] = 0.5 * eta[3 * q + c] + 0.5 * eta[3 * r + c];
        fzet[3 * q + c] = 0.5 * zet[3 * q + c] + 0.5 * zet[3 * r + c];
    }
    faj[q] = 2. / (1. / aj[q] + 1. / aj[r]);
}
#undef CRD

int launch_metrics(vfs_ctx *c, const Grid &g, const double *coor, double *const m[17])
{
    hipStream_t st = c->stream;
    const size_t n = (size_t)g.n;
    for (int q = 0; q < 17; q++) {
        bool vec = !(q == 3 || q == 7 || q == 11 || q == 15);
        VFS_HIP(hipMemsetAsync(m[q], 0, (vec ? 3 : 1) * n * sizeof(double), st));
    }
    const unsigned nb = blocks_for(g.n, TPB);
    hipLaunchKernelGGL(k_metrics_cell, dim3(nb), dim3(TPB), 0, st, g, coor, m[0], m[1], m[2], m[3], m[16]);
    VFS_LAUNCH_CHECK("metrics cell");
    const long long pl[3] = {(long long)g.my * g.mz, (long long)g.mx * g.mz, (long long)g.mx * g.my};
    for (int w = 0; w < 6; w++)
        hipLaunchKernelGGL(k_metrics_mirror, dim3(blocks_for(pl[w >> 1], TPB)), dim3(TPB), 0, st, g, w, m[0], m[1], m[2], m[3]);
    VFS_LAUNCH_CHECK("metrics mirror");
    // periodic copy into the cell arrays before the face means read them
    // (metrics.c:365-395)
    for (int q = 0; q < 4; q++) {
        int rc = periodic_fix_grid(st, g, m[q], q == 3 ? 1 : 3, 0);
        if (rc) return rc;
    }
    for (int axis = 0; axis < 3; axis++)
        hipLaunchKernelGGL(k_face_metrics, dim3(nb), dim3(TPB), 0, st, g, axis, m[0], m[1], m[2], m[3],
                           m[4 + 4 * axis], m[5 + 4 * axis], m[6 + 4 * axis], m[7 + 4 * axis]);
    VFS_LAUNCH_CHECK("face metrics");
    // final periodic fix of the sixteen arrays (metrics.c:1233-1280)
    for (int q = 0; q < 16; q++) {
        bool vec = !(q == 3 || q == 7 || q == 11 || q == 15);
        int rc = periodic_fix_grid(st, g, m[q], vec ? 3 : 1, 0);
        if (rc) return rc;
    }
    return VFS_OK;
}

// ---------------------------------------------------------------------------
// a2: Contra2Cart_2 (rhs.c:65-749)
// ---------------------------------------------------------------------------

// interior Cramer solve, rhs.c:158-247
__device__ inline void c2c_point(const Grid &g, int k, int j, int i, const double *__restrict__ ucont,
                                 const double *__restrict__ csi, const double *__restrict__ eta,
                                 const double *__restrict__ zet, const double *__restrict__ nvert,
                                 double *__restrict__ ucat)
{
    const long long q = gidx(g, k, j, i);
    if (nvert && nvert[q] >= 0.1) return;  // nvert == nullptr: no point is blanked
    double q0 = 0.5 * (ucont[3 * gidx(g, k, j, i - 1) + 0] + ucont[3 * q + 0]);
    double q1 = 0.5 * (ucont[3 * gidx(g, k, j - 1, i) + 1] + ucont[3 * q + 1]);
    double q2 = 0.5 * (ucont[3 * gidx(g, k - 1, j, i) + 2] + ucont[3 * q + 2]);
    const double m00 = csi[3 * q], m01 = csi[3 * q + 1], m02 = csi[3 * q + 2];
    const double m10 = eta[3 * q], m11 = eta[3 * q + 1], m12 = eta[3 * q + 2];
    const double m20 = zet[3 * q], m21 = zet[3 * q + 1], m22 = zet[3 * q + 2];
    double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
    double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
    double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
    ucat[3 * q + 0] = det0 / det;
    ucat[3 * q + 1] = det1 / det;
    ucat[3 * q + 2] = det2 / det;
}

__global__ void k_c2c_interior(Grid g, const double *__restrict__ ucont, const double *__restrict__ csi,
                               const double *__restrict__ eta, const double *__restrict__ zet,
                               const double *__restrict__ nvert, double *__restrict__ ucat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1) return;
    c2c_point(g, k, j, i, ucont, csi, eta, zet, nvert, ucat);
}

// the interior cells whose Contra2Cart inputs k_wall_flux may change: the
// planes i = 1 / i = mx-2 / j = 1 / j = my-2 next to a wall face (bc 1 or 10)
__device__ inline bool wall_face(const Grid &g, int side) { return g.bc[side] == 1 || g.bc[side] == 10; }
__global__ void k_c2c_walls(Grid g, const double *__restrict__ ucont, const double *__restrict__ csi,
                            const double *__restrict__ eta, const double *__restrict__ zet,
                            const double *__restrict__ nvert, double *__restrict__ ucat)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ni = (long long)(g.my - 2) * (g.mz - 2), nj = (long long)(g.mx - 2) * (g.mz - 2);
    const int side = (int)(t / (ni > nj ? ni : nj));
    const long long r = t % (ni > nj ? ni : nj);
    if (side >= 4 || !wall_face(g, side)) return;
    int i, j, k;
    if (side < 2) {
        if (r >= ni) return;
        j = 1 + (int)(r % (g.my - 2));
        k = 1 + (int)(r / (g.my - 2));
        i = side == 0 ? 1 : g.mx - 2;
    } else {
        if (r >= nj) return;
        i = 1 + (int)(r % (g.mx - 2));
        k = 1 + (int)(r / (g.mx - 2));
        j = side == 2 ? 1 : g.my - 2;
        // cells of an i-wall plane are done by that plane's threads
        if ((i == 1 && wall_face(g, 0)) || (i == g.mx - 2 && wall_face(g, 1))) return;
    }
    c2c_point(g, k, j, i, ucont, csi, eta, zet, nvert, ucat);
}

// slip reflection, rhs.c:447-544
__device__ inline void slip_fill(const Grid &g, const double *csi, const double *eta, const double *zet,
                                 const double *uold, int kn, int jn, int in, int dir, double sgn, double *out)
{
    long long q = gidx(g, kn, jn, in);
    double gm[3][3], G[3][3];
    for (int c = 0; c < 3; c++) { gm[0][c] = csi[3 * q + c]; gm[1][c] = eta[3 * q + c]; gm[2][c] = zet[3 * q + c]; }
    covariant(gm, G);
    double nx = sgn * G[0][dir], ny = sgn * G[1][dir], nz = sgn * G[2][dir];
    double sum = sqrt(nx * nx + ny * ny + nz * nz);
    nx /= sum; ny /= sum; nz /= sum;
    double ux = uold[3 * q], uy = uold[3 * q + 1], uz = uold[3 * q + 2];
    double un = ux * nx + uy * ny + uz * nz;
    out[0] = ux - 2 * un * nx;
    out[1] = uy - 2 * un * ny;
    out[2] = uz - 2 * un * nz;
}

// ghost-cell BC fill, rhs.c:300-680, for the boundary planes of the
// non-periodic directions.  Each ghost cell is visited by exactly one thread
// (planes of a later axis skip cells owned by an earlier axis).
__global__ void k_c2c_bc(Grid g, int axis, const double *__restrict__ L2, const double *__restrict__ uold,
                         const double *__restrict__ csi, const double *__restrict__ eta,
                         const double *__restrict__ zet, const double *__restrict__ nvert, double *__restrict__ ucat)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (axis == 0) {
        if (t >= 2LL * g.my * g.mz) return;
        int side = t & 1; long long r = t >> 1;
        j = r % g.my; k = r / g.my; i = side ? g.mx - 1 : 0;
    } else if (axis == 1) {
        if (t >= 2LL * g.mx * g.mz) return;
        i = t % g.mx; long long r = t / g.mx;
        int side = r & 1; k = r >> 1; j = side ? g.my - 1 : 0;
        if (!g.pi && (i == 0 || i == g.mx - 1)) return;
    } else {
        if (t >= 2LL * g.mx * g.my) return;
        i = t % g.mx; long long r = t / g.mx;
        j = r % g.my; int side = (int)(r / g.my); k = side ? g.mz - 1 : 0;
        if (!g.pi && (i == 0 || i == g.mx - 1)) return;
        if (!g.pj && (j == 0 || j == g.my - 1)) return;
    }
    const int *bc = g.bc;
    const long long q = gidx(g, k, j, i);
    double u[3] = {ucat[3 * q], ucat[3 * q + 1], ucat[3 * q + 2]};
    bool touched = false;
    int solid_flag = 0;
    const int mx = g.mx, my = g.my, mz = g.mz;
    if (bc[0] == 10 && i == 0 && (j != 0 && k != 0)) {
        slip_fill(g, csi, eta, zet, uold, k, j, i + 1, 0, -1.0, u);
        if (nvert[gidx(g, k, j, i + 1)] > 0.1) u[0] = u[1] = u[2] = 0;
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    if (bc[1] == 10 && i == mx - 1 && (j != 0 && k != 0)) {
        slip_fill(g, csi, eta, zet, uold, k, j, i - 1, 0, 1.0, u);
        if (nvert[gidx(g, k, j, i - 1)] > 0.1) u[0] = u[1] = u[2] = 0;
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    if (bc[2] == 10 && j == 0 && (i != 0 && k != 0)) {
        slip_fill(g, csi, eta, zet, uold, k, j + 1, i, 1, -1.0, u);
        if (nvert[gidx(g, k, j + 1, i)] > 0.1) u[0] = u[1] = u[2] = 0;
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    if ((bc[3] == 10 || bc[3] == -10) && j == my - 1 && (i != 0 && k != 0)) {
        slip_fill(g, csi, eta, zet, uold, k, j - 1, i, 1, 1.0, u);
        if (nvert[gidx(g, k, j - 1, i)] > 0.1) u[0] = u[1] = u[2] = 0;
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    if (i == 0 && bc[0] == 1 && (j != 0 && k != 0)) {
        long long s = gidx(g, k, j, i + 1);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    if (i == mx - 1 && bc[1] == 1 && (j != 0 && k != 0)) {
        long long s = gidx(g, k, j, i - 1);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    if (j == 0 && bc[2] == 1 && (i != 0 && k != 0)) {
        long long s = gidx(g, k, j + 1, i);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    if (j == my - 1 && bc[3] == 1 && (i != 0 && k != 0)) {
        long long s = gidx(g, k, j - 1, i);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    if (k == 0 && bc[4] == 1 && (i != 0 && j != 0)) {
        long long s = gidx(g, k + 1, j, i);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    if (k == mz - 1 && bc[5] == 1 && (i != 0 && j != 0)) {
        long long s = gidx(g, k - 1, j, i);
        for (int c = 0; c < 3; c++) u[c] = -L2[3 * s + c];
        solid_flag = 1; touched = true;
    }
    // lid-driven cavity j+ (rhs.c:593-605): the wall moves with u = 1
    if (j == my - 1 && bc[3] == 2) {
        const double *nb = (solid_flag ? uold : L2) + 3 * gidx(g, k, j - 1, i);
        u[0] = 2.0 - nb[0]; u[1] = -nb[1]; u[2] = -nb[2];
        touched = true;
    }
    // Couette walls j- / j+ (rhs.c:608-624): the wall moves with w = 1
    if (j == 0 && bc[2] == 12) {
        const double *nb = L2 + 3 * gidx(g, k, j + 1, i);
        u[0] = -nb[0]; u[1] = -nb[1]; u[2] = 2.0 - nb[2];
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    if (j == my - 1 && bc[3] == 12) {
        const double *nb = L2 + 3 * gidx(g, k, j - 1, i);
        u[0] = -nb[0]; u[1] = -nb[1]; u[2] = 2.0 - nb[2];
        if (solid_flag) u[0] = u[1] = u[2] = 0;
        touched = true;
    }
    // outflow j+ (rhs.c:640-642)
    if (bc[3] == 4 && j == my - 1 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1)) {
        const long long s = gidx(g, k, j - 1, i);
        for (int c = 0; c < 3; c++) u[c] = L2[3 * s + c];
        touched = true;
    }
    if (bc[5] == 4 && k == mz - 1 && (i != 0 && i != mx - 1 && j != 0 && j != my - 1)) {
        if (nvert[gidx(g, k - 1, j, i)] > 0.1) { u[0] = u[1] = u[2] = 0; touched = true; }
        if (solid_flag) { u[0] = u[1] = u[2] = 0; touched = true; }
    }
    if (touched) { ucat[3 * q] = u[0]; ucat[3 * q + 1] = u[1]; ucat[3 * q + 2] = u[2]; }
}

// cells next to two walls (rhs.c:674-680): every k but 0 and mz-2, the k+
// ghost plane included (the reference's loop spans the whole DA)
__global__ void k_c2c_corner(Grid g, double *__restrict__ ucat)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 1 || k >= g.mz || k == g.mz - 2) return;
    if (g.my - 2 == 1 || g.mx - 1 == 1) return;
    const int *bc = g.bc;
    if (bc[0] <= 1 && bc[2] <= 1) { long long q = gidx(g, k, 1, 1); ucat[3 * q] = ucat[3 * q + 1] = ucat[3 * q + 2] = 0; }
    if (bc[1] <= 1 && bc[2] <= 1) { long long q = gidx(g, k, 1, g.mx - 2); ucat[3 * q] = ucat[3 * q + 1] = ucat[3 * q + 2] = 0; }
}

static bool bc_needs_snapshot(const Grid &g)
{
    // the cavity, Couette and j+ outflow rules read the pre-pass values of cells
    // another axis' rule may write
    if (g.bc[3] == 2 || g.bc[2] == 12 || g.bc[3] == 12 || g.bc[3] == 4) return true;
    // a no-slip rule reads the neighbouring plane, which another axis' rule
    // may write in the same pass only if walls sit on two or more axes
    const int per[3] = {g.pi, g.pj, g.pk};
    int axes_noslip = 0, axes_write = 0;
    for (int a = 0; a < 3; a++) {
        if (per[a]) continue;
        bool ns = g.bc[2 * a] == 1 || g.bc[2 * a + 1] == 1;
        bool wr = ns || g.bc[2 * a] == 10 || g.bc[2 * a + 1] == 10 || (a == 1 && g.bc[3] == -10) ||
                  (a == 2 && g.bc[5] == 4);
        axes_noslip += ns;
        axes_write += wr;
    }
    return axes_noslip >= 1 && axes_write >= 2;
}

static bool corner_rule_active(const Grid &g)
{
    return (g.bc[0] <= 1 && g.bc[2] <= 1) || (g.bc[1] <= 1 && g.bc[2] <= 1);
}

int c2c(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    if (c->ucat_fresh) return VFS_OK;  // would rewrite the same values
    double *ucont = c->f[VFS_UCONT], *ucat = c->f[VFS_UCAT];
    int rc = periodic_fix(c, ucont, 3, 0);  // rhs.c:129-156
    if (rc) return rc;
    // The interior cells are a pointwise function of Ucont, the metrics and
    // nvert; where they are known to hold that function already, only the
    // cells whose inputs changed are recomputed.  The ghost-cell passes below
    // then see the same interior and the same ghost values as after a full
    // interior pass.
    const char *pe = std::getenv("VFS_C2C_PARTIAL");
    const int part = pe && pe[0] == '0' ? (int)UCAT_STALE : c->ucat_part;
    bool nb = false;
    rc = nvert_all_zero(c, &nb);
    if (rc) return rc;
    const double *nvi = nb ? nullptr : c->f[VFS_NVERT];  // interior passes skip the blanking test
    if (part == UCAT_WALLS) {
        const long long ni = (long long)(g.my - 2) * (g.mz - 2), nj = (long long)(g.mx - 2) * (g.mz - 2);
        hipLaunchKernelGGL(k_c2c_walls, dim3(blocks_for(4 * (ni > nj ? ni : nj), TPB)), dim3(TPB), 0, st, g, ucont,
                           c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], nvi, ucat);
        VFS_LAUNCH_CHECK("c2c walls");
    } else if (part != UCAT_INTERIOR) {
        dim3 blk(TPB), grd(blocks_for(g.mx - 2, TPB), g.my - 2, g.mz - 2);
        prof_begin(c, K_C2C);
        hipLaunchKernelGGL(k_c2c_interior, grd, blk, 0, st, g, ucont, c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], nvi,
                           ucat);
        prof_end(c);
        VFS_LAUNCH_CHECK("c2c interior");
    }
    rc = periodic_fix(c, ucat, 3, 0);  // rhs.c:254-291
    if (rc) return rc;
    const double *L2 = ucat;
    const bool corner = corner_rule_active(g);
    if (bc_needs_snapshot(g) || corner) {
        VFS_HIP(hipMemcpyAsync(c->snap, ucat, 3 * (size_t)g.n * sizeof(double), hipMemcpyDeviceToDevice, st));
        L2 = c->snap;
    }
    const int per[3] = {g.pi, g.pj, g.pk};
    const long long np[3] = {2LL * g.my * g.mz, 2LL * g.mx * g.mz, 2LL * g.mx * g.my};
    for (int a = 0; a < 3; a++) {
        if (per[a]) continue;
        hipLaunchKernelGGL(k_c2c_bc, dim3(blocks_for(np[a], TPB)), dim3(TPB), 0, st, g, a, L2, c->f[VFS_UCAT_OLD],
                           c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_NVERT], ucat);
    }
    VFS_LAUNCH_CHECK("c2c bc");
    if (corner) {
        hipLaunchKernelGGL(k_c2c_corner, dim3(blocks_for(g.mz, TPB)), dim3(TPB), 0, st, g, ucat);
        VFS_LAUNCH_CHECK("c2c corner");
    }
    rc = periodic_fix(c, ucat, 3, 0);  // rhs.c:712-748
    if (!rc) {
        c->ucat_fresh = true;
        c->ucat_part = UCAT_STALE;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// FormBCS subset (bcs.c:1684-2767)
// ---------------------------------------------------------------------------

// outlet plane sums: FluxOut and the area (bcs.c:2127-2175)
__global__ void k_outlet_sums(Grid g, const double *__restrict__ ucont, const double *__restrict__ zet,
                              const double *__restrict__ nvert, double *__restrict__ part)
{
    const int k = g.mz - 1;
    const long long n = (long long)(g.mx - 2) * (g.my - 2);
    double flux = 0, area = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i = 1 + t % (g.mx - 2), j = 1 + t / (g.mx - 2);
        long long q = gidx(g, k - 1, j, i);
        if (nvert[q] < 0.1) {
            flux += ucont[3 * q + 2];
            area += sqrt(zet[3 * q] * zet[3 * q] + zet[3 * q + 1] * zet[3 * q + 1] + zet[3 * q + 2] * zet[3 * q + 2]);
        }
    }
    flux = block_reduce<OpSum>(flux);
    area = block_reduce<OpSum>(area);
    if (threadIdx.x == 0) { part[blockIdx.x] = flux; part[RED_BLOCKS + blockIdx.x] = area; }
}

__global__ void k_outlet_scale(Grid g, double flux_in, const double *__restrict__ sums, const double *__restrict__ zet,
                               const double *__restrict__ nvert, double *__restrict__ ucont)
{
    const int k = g.mz - 1;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)(g.mx - 2) * (g.my - 2)) return;
    int i = 1 + t % (g.mx - 2), j = 1 + t / (g.mx - 2);
    long long q = gidx(g, k - 1, j, i);
    double Area = sqrt(zet[3 * q] * zet[3 * q] + zet[3 * q + 1] * zet[3 * q + 1] + zet[3 * q + 2] * zet[3 * q + 2]);
    if (nvert[q] < 0.1) ucont[3 * q + 2] += (flux_in - sums[0]) * Area / sums[1];
}

// wall-normal flux zeroing, bcs.c:2347-2386
__global__ void k_wall_flux(Grid g, double *__restrict__ ucont)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int *bc = g.bc;
    // i faces: (k,j) interior
    long long ni = (long long)(g.my - 2) * (g.mz - 2);
    if (t < ni) {
        int j = 1 + t % (g.my - 2), k = 1 + t / (g.my - 2);
        if (bc[0] == 1 || bc[0] == 10) ucont[3 * gidx(g, k, j, 0) + 0] = 0;
        if (bc[1] == 1 || bc[1] == 10) ucont[3 * gidx(g, k, j, g.mx - 2) + 0] = 0;
    }
    long long nj = (long long)(g.mx - 2) * (g.mz - 2);
    if (t < nj) {
        int i = 1 + t % (g.mx - 2), k = 1 + t / (g.mx - 2);
        if (bc[2] == 1 || bc[2] == 10) ucont[3 * gidx(g, k, 0, i) + 1] = 0;
        if (bc[3] == 1 || bc[3] == 10) ucont[3 * gidx(g, k, g.my - 2, i) + 1] = 0;
    }
}

// zero the twelve ghost edges of Ucat, bcs.c:2630-2760
__global__ void k_zero_edges(Grid g, double *__restrict__ ucat)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    auto z = [&](int k, int j, int i) {
        long long q = gidx(g, k, j, i);
        ucat[3 * q] = ucat[3 * q + 1] = ucat[3 * q + 2] = 0;
    };
    if (t < g.my) { int j = (int)t; z(0, j, 0); z(0, j, g.mx - 1); z(g.mz - 1, j, 0); z(g.mz - 1, j, g.mx - 1); }
    if (t < g.mx) { int i = (int)t; z(0, 0, i); z(0, g.my - 1, i); z(g.mz - 1, 0, i); z(g.mz - 1, g.my - 1, i); }
    if (t < g.mz) { int k = (int)t; z(k, 0, 0); z(k, 0, g.mx - 1); z(k, g.my - 1, 0); z(k, g.my - 1, g.mx - 1); }
}

// ---------------------------------------------------------------------------
// InflowFlux (bcs.c:247-626): the inlet on the k- face
// ---------------------------------------------------------------------------

// InletInterpolation (bcs.c:636-681) returns main.c:13's constant Flux_in
// before reaching its table code
constexpr double FLUX_IN_CONST = 4.104388e-04;

// uin of bcs.c:308-371 at an inlet face with centre (xc, yc), radius r;
// uniform = 1 or -flux / inletArea (profiles 1 and 2)
__device__ inline double inlet_uin(int profile, double uniform, double yc, double r)
{
    switch (profile) {
    case 0:
    case 3: return FLUX_IN_CONST;
    case 10: {
        const double delta = 0.45263, a = 5.99;
        if (yc >= delta) return 1.0;
        if (yc <= 0) return 0.0;
        return pow(yc / delta, 1. / a);
    }
    case 12: return 2 * (1 - pow(r / 0.5, 2.0));
    case 14: {
        const double R[11] = {0.000, 0.062, 0.125, 0.188, 0.250, 0.312, 0.354, 0.399, 0.438, 0.469, 0.500};
        const double W[11] = {1.120, 1.115, 1.110, 1.093, 1.050, 1.010, 0.963, 0.915, 0.825, 0.730, 0.000};
        int ii;
        for (ii = 1; ii < 11; ii++)
            if (R[ii] >= r && R[ii - 1] < r) break;
        if (ii == 11) ii = 10;  // r = 0 or r > 0.5 (replaced by 0 below); the reference reads past its table
        const double u = (W[ii] - W[ii - 1]) / (R[ii] - R[ii - 1]) * (r - R[ii - 1]) + W[ii - 1];
        return r > 0.5 ? 0 : u;
    }
    case 15: return r > 0.5 ? 0 : 1;
    case 21: return 1.5 * 1.0 * yc * (2 - yc);
    default: return uniform;  // 1, 2 (100 uses the plane)
    }
}

// mode 0: partial sums of inletArea (bcs.c:201-234);
// mode 1: the inlet fluxes and ghost velocities (bcs.c:379-449) and partial
//         sums of the inlet flux;
// mode 2: the held-flux correction ucont.z += (flux0 - flux1) |zet| / inletArea
//         (bcs.c:487-496, sums[0] = flux1, sums[1] = inletArea) and partial
//         sums of the corrected fluid inlet flux
__global__ void k_inflow(Grid g, int mode, int profile, int plane, double inlet_flux, double flux0,
                         const double *__restrict__ sums, const double *__restrict__ coor,
                         const double *__restrict__ pl, const double *__restrict__ csi, const double *__restrict__ eta,
                         const double *__restrict__ zet, const double *__restrict__ nvert, double *__restrict__ ucont,
                         double *__restrict__ ucat, double *__restrict__ part)
{
    const long long n = (long long)(g.mx - 2) * (g.my - 2);
    double s = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + t % (g.mx - 2), j = 1 + t / (g.mx - 2);
        const long long q = gidx(g, 0, j, i), q1 = gidx(g, 1, j, i);
        const double *z = zet + 3 * q;
        const double area = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
        const bool fluid = nvert[q1] < 0.1;
        if (mode == 0) {
            if (fluid) s += area;
            continue;
        }
        if (mode == 2) {
            if (fluid) {
                ucont[3 * q + 2] += (flux0 - sums[0]) * area / sums[1];
                s += ucont[3 * q + 2];
            }
            continue;
        }
        if (fluid) {
            if (plane) {
                const double *u = pl + 3 * ((long long)j * g.mx + i);
                for (int c = 0; c < 3; c++) ucat[3 * q + c] = u[c];
                ucont[3 * q + 2] = 0.5 * (ucat[3 * q] + ucat[3 * q1]) * z[0] + 0.5 * (ucat[3 * q + 1] + ucat[3 * q1 + 1]) * z[1] +
                                   0.5 * (ucat[3 * q + 2] + ucat[3 * q1 + 2]) * z[2];
            } else {
                // face centre of the inlet cell from the k = 1 node plane (bcs.c:307-308)
                const long long a0 = gidx(g, 1, j, i), a1 = gidx(g, 1, j - 1, i), a2 = gidx(g, 1, j, i - 1),
                                a3 = gidx(g, 1, j - 1, i - 1);
                const double xc = (coor[3 * a0] + coor[3 * a1] + coor[3 * a2] + coor[3 * a3]) * 0.25;
                const double yc = (coor[3 * a0 + 1] + coor[3 * a1 + 1] + coor[3 * a2 + 1] + coor[3 * a3 + 1]) * 0.25;
                const double r = sqrt(xc * xc + yc * yc);
                const double uniform = inlet_flux < 0 ? 1. : inlet_flux / sums[1];
                const double q2 = inlet_uin(profile, uniform, yc, r) * area;
                ucont[3 * q + 2] = q2;
                ucont[3 * q] = 0;
                ucont[3 * q + 1] = 0;
                // Contra2Cart_single (rhs.c:42-63) of (0, 0, q2)
                const double m00 = csi[3 * q], m01 = csi[3 * q + 1], m02 = csi[3 * q + 2];
                const double m10 = eta[3 * q], m11 = eta[3 * q + 1], m12 = eta[3 * q + 2];
                const double m20 = z[0], m21 = z[1], m22 = z[2];
                const double q0 = 0, qq1 = 0;
                const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
                const double det0 = q0 * (m11 * m22 - m12 * m21) - qq1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
                const double det1 = -q0 * (m10 * m22 - m12 * m20) + qq1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
                const double det2 = q0 * (m10 * m21 - m11 * m20) - qq1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
                const double u[3] = {det0 / det, det1 / det, det2 / det};
                for (int c = 0; c < 3; c++) ucat[3 * q + c] = -ucat[3 * q1 + c] + 2 * u[c];
            }
            s += ucont[3 * q + 2];
        } else {
            ucat[3 * q + 2] = 0;
            ucont[3 * q + 2] = 0;
        }
    }
    s = block_reduce<OpSum>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// pseudo_periodic_BC (bcs.c:29-95): plane := ucat(kplane), kplane = (6 mz) / 10
__global__ void k_pseudo_plane(Grid g, const double *__restrict__ ucat, double *__restrict__ pl)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= 3 * g.plane) return;
    const long long kplane = (g.mz * 6) / 10;
    pl[t] = ucat[3 * kplane * g.plane + t];
}

int pseudo_periodic_plane(vfs_ctx *c)
{
    const long long n = 3 * c->g.plane;
    hipLaunchKernelGGL(k_pseudo_plane, dim3(blocks_for(n, TPB)), dim3(TPB), 0, c->stream, c->g, c->f[VFS_UCAT], c->inflow_plane);
    VFS_LAUNCH_CHECK("pseudo-periodic plane");
    return VFS_OK;
}

int inflow_flux(vfs_ctx *c, double *flux_in)
{
    const Grid &g = c->g;
    if (g.bc[4] != 5) {
        if (flux_in) *flux_in = c->p.flux_in_sum;
        return VFS_OK;
    }
    hipStream_t st = c->stream;
    const int profile = c->p.inletprofile;
    const int plane = c->p.pseudo_periodic || profile == 100;
    const bool held = c->p.pseudo_periodic || profile == 100 || profile == 14;
    const double inlet_flux = c->p.inlet_flux;
    // device scalars: sums[0] = flux of the set inlet (flux1), sums[1] = inletArea, sums[2] = FluxIn
    double *sums = c->red + 2 * RED_BLOCKS;
    hipLaunchKernelGGL(k_inflow, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, 0, profile, plane, inlet_flux, 0.0,
                       (const double *)nullptr, c->f[VFS_COOR], c->inflow_plane, c->f[VFS_CSI], c->f[VFS_ETA],
                       c->f[VFS_ZET], c->f[VFS_NVERT], c->f[VFS_UCONT], c->f[VFS_UCAT], c->red);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, sums + 1);
    hipLaunchKernelGGL(k_inflow, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, 1, profile, plane, inlet_flux, 0.0,
                       (const double *)sums, c->f[VFS_COOR], c->inflow_plane, c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET],
                       c->f[VFS_NVERT], c->f[VFS_UCONT], c->f[VFS_UCAT], c->red);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, sums);
    double *result = sums;
    if (held && inlet_flux > 0) {
        hipLaunchKernelGGL(k_inflow, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, 2, profile, plane, inlet_flux, inlet_flux,
                           (const double *)sums, c->f[VFS_COOR], c->inflow_plane, c->f[VFS_CSI], c->f[VFS_ETA],
                           c->f[VFS_ZET], c->f[VFS_NVERT], c->f[VFS_UCONT], c->f[VFS_UCAT], c->red);
        hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, sums + 2);
        result = sums + 2;
    }
    ucat_stale(c);
    VFS_LAUNCH_CHECK("inflow flux");
    VFS_HIP(hipMemcpyAsync(c->red_host, result, sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    c->p.flux_in_sum = c->red_host[0];  // FluxInSum, read by the outlet balance of FormBCS
    if (flux_in) *flux_in = c->red_host[0];
    return VFS_OK;
}

// Boundary-face loops of FormBCS.  Face f (0..5 = i-, i+, j-, j+, k-, k+) is
// iterated over the interior of the other two directions; (k, j, i) of its
// t-th point on the ghost layer.
__device__ inline bool face_point(const Grid &g, int f, long long t, int &k, int &j, int &i)
{
    const int a = f / 2;
    const int n0 = a == 0 ? g.my - 2 : g.mx - 2, n1 = a == 2 ? g.my - 2 : g.mz - 2;
    if (t >= (long long)n0 * n1) return false;
    const int u = 1 + (int)(t % n0), v = 1 + (int)(t / n0);
    const int w = (f & 1) ? (a == 0 ? g.mx - 1 : a == 1 ? g.my - 1 : g.mz - 1) : 0;
    if (a == 0) { i = w; j = u; k = v; }
    else if (a == 1) { j = w; i = u; k = v; }
    else { k = w; i = u; j = v; }
    return true;
}

// far field (code 6), bcs.c:1785-1888: the boundary flux of face f from the
// adjacent cell's velocity component along the face's own axis times the
// cell's matching metric component; partial sums of the flux and of the
// ghost-face area (in: faces i-, j-, k-; out: i+, j+, k+).  With mode 1 the
// flux-difference correction of bcs.c:1890-1993 (dv[0] = VelDiffIn,
// dv[1] = VelDiffOut) is applied instead.
__global__ void k_farfield(Grid g, int f, int mode, const double *__restrict__ dv, const double *__restrict__ ucat,
                           const double *__restrict__ m_axis, double *__restrict__ ucont, double *__restrict__ part)
{
    const int a = f / 2, hi = f & 1;
    double flux = 0, area = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;; t += (long long)gridDim.x * blockDim.x) {
        int k, j, i;
        if (!face_point(g, f, t, k, j, i)) break;
        // the face stored at the ghost index (low side) or at the last interior index (high side)
        int kf = k, jf = j, if_ = i, kc = k, jc = j, ic = i;
        if (hi) { if (a == 0) if_ = ic = i - 1; else if (a == 1) jf = jc = j - 1; else kf = kc = k - 1; }
        else { if (a == 0) ic = i + 1; else if (a == 1) jc = j + 1; else kc = k + 1; }
        const long long qf = gidx(g, kf, jf, if_), qc = gidx(g, kc, jc, ic), qg = gidx(g, k, j, i);
        double v = ucat[3 * qc + a];
        if (mode) v = hi ? v + dv[1] : v - dv[0];
        ucont[3 * qf + a] = v * m_axis[3 * qc + a];
        flux += ucont[3 * qf + a];
        area += m_axis[3 * (hi ? qf : qg) + a];
    }
    if (mode) return;
    flux = block_reduce<OpSum>(flux);
    area = block_reduce<OpSum>(area);
    if (threadIdx.x == 0) { part[blockIdx.x] = flux; part[RED_BLOCKS + blockIdx.x] = area; }
}

// VelDiffIn / VelDiffOut (bcs.c:1890-1902) from s = {FluxIn, AreaIn, FluxOut, AreaOut}
__global__ void k_farfield_diff(const double *__restrict__ s, double *__restrict__ dv)
{
    const double FluxDiff = 0.5 * (s[0] - s[2]);
    double in = FluxDiff / s[1], out = FluxDiff / s[3];
    if (fabs(FluxDiff) < 1.e-6 || fabs(s[1]) < 1.e-6) in = 0.;
    if (fabs(FluxDiff) < 1.e-6 || fabs(s[3]) < 1.e-6) out = 0.;
    dv[0] = in;
    dv[1] = out;
}

// partial sums over the interior of the k- ghost plane (mode 0: k- outlet,
// bcs.c:2294-2320, the face flux of the first cell's velocity and zet.z of
// fluid ghost cells) or over the faces k = mz-2 (mode 1: the characteristic
// outlet's FluxOut, bcs.c:1997-2012)
__global__ void k_kplane_sums(Grid g, int mode, const double *__restrict__ ucat, const double *__restrict__ ucont,
                              const double *__restrict__ zet, const double *__restrict__ nvert, double *__restrict__ part)
{
    const long long n = (long long)(g.mx - 2) * (g.my - 2);
    double s0 = 0, s1 = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + t % (g.mx - 2), j = 1 + t / (g.mx - 2);
        if (mode == 1) {
            s0 += ucont[3 * gidx(g, g.mz - 2, j, i) + 2];
            continue;
        }
        const long long q0 = gidx(g, 0, j, i), q1 = gidx(g, 1, j, i);
        if (nvert[q0] < 0.1) {
            s0 += (ucat[3 * q1] * (zet[3 * q0] + zet[3 * q1]) + ucat[3 * q1 + 1] * (zet[3 * q0 + 1] + zet[3 * q1 + 1]) +
                   ucat[3 * q1 + 2] * (zet[3 * q0 + 2] + zet[3 * q1 + 2])) * 0.5;
            s1 += zet[3 * q0 + 2];
        }
    }
    s0 = block_reduce<OpSum>(s0);
    s1 = block_reduce<OpSum>(s1);
    if (threadIdx.x == 0) { part[blockIdx.x] = s0; part[RED_BLOCKS + blockIdx.x] = s1; }
}

// mode 0: the k- outlet fluxes ucont(0).z = (ucat(1).z + ratio) zet(0).z with
// ratio = (FluxInSum - s[0]) / s[1] (bcs.c:2322-2338); mode 1: the
// characteristic outlet's scaling ucont(mz-2).z *= ratio with ratio =
// (FluxInSum + FarFluxIn) / FluxOut (bcs.c:2013-2037; s[0] = FluxOut, s[2] = FarFluxIn)
__global__ void k_kplane_set(Grid g, int mode, double flux_in_sum, const double *__restrict__ s,
                             const double *__restrict__ ucat, const double *__restrict__ zet, double *__restrict__ ucont)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)(g.mx - 2) * (g.my - 2)) return;
    const int i = 1 + t % (g.mx - 2), j = 1 + t / (g.mx - 2);
    if (mode == 0) {
        const double ratio = (flux_in_sum - s[0]) / s[1];
        const long long q0 = gidx(g, 0, j, i), q1 = gidx(g, 1, j, i);
        ucont[3 * q0 + 2] = (ucat[3 * q1 + 2] + ratio) * zet[3 * q0 + 2];
    } else {
        const double FluxIn = flux_in_sum + s[2];
        double ratio = FluxIn / s[0];
        if (fabs(s[0]) < 1.e-6) ratio = 1.;
        if (fabs(FluxIn) < 1.e-6) ratio = 0.;
        const long long q = gidx(g, g.mz - 2, j, i);
        ucont[3 * q + 2] = ucont[3 * q + 2] * ratio;
    }
}

// far field and characteristic outlet of FormBCS (bcs.c:1745-2041);
// scal[0..3] = far-field sums, scal[4..5] = VelDiffIn/Out, scal[6] = the
// characteristic outlet's FluxOut, scal[8] = FarFluxIn again
static int form_bcs_open(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    double *part = c->red, *scal = c->red + 3 * RED_BLOCKS;
    const double *m_axis[3] = {c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET]};
    bool far = false, changed = false;
    for (int f = 0; f < 6; f++) far = far || g.bc[f] == 6;
    VFS_HIP(hipMemsetAsync(scal, 0, 9 * sizeof(double), st));
    if (far) {
        for (int f = 0; f < 6; f++) {
            if (g.bc[f] != 6) continue;
            hipLaunchKernelGGL(k_farfield, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, f, 0, (const double *)nullptr,
                               c->f[VFS_UCAT], m_axis[f / 2], c->f[VFS_UCONT], part);
            // accumulate into scal[0..1] (in) or scal[2..3] (out) in face order
            const int o = (f & 1) ? 2 : 0;
            hipLaunchKernelGGL(finish_reduce_add<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, part, RED_BLOCKS, scal + o);
            hipLaunchKernelGGL(finish_reduce_add<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, part + RED_BLOCKS, RED_BLOCKS,
                               scal + o + 1);
        }
        if (g.bc[5] == 6) {
            hipLaunchKernelGGL(k_farfield_diff, dim3(1), dim3(1), 0, st, scal, scal + 4);
            const int order[6] = {5, 3, 1, 0, 2, 4};  // bcs.c:1922-1993
            for (int f : order) {
                if (g.bc[f] != 6) continue;
                hipLaunchKernelGGL(k_farfield, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, f, 1, scal + 4,
                                   c->f[VFS_UCAT], m_axis[f / 2], c->f[VFS_UCONT], part);
            }
        }
        changed = true;
    }
    const long long np = (long long)(g.mx - 2) * (g.my - 2);
    if (g.bc[5] == 8) {
        hipLaunchKernelGGL(k_kplane_sums, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, 1, c->f[VFS_UCAT], c->f[VFS_UCONT],
                           c->f[VFS_ZET], c->f[VFS_NVERT], part);
        hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, part, RED_BLOCKS, scal + 6);
        // k_kplane_set mode 1 reads FluxOut at s[0] and FarFluxIn at s[2]
        VFS_HIP(hipMemcpyAsync(scal + 8, scal, sizeof(double), hipMemcpyDeviceToDevice, st));
        // the scaling starts at the third step (bcs.c:2026-2036)
        if (!(c->ti == 0 || c->ti == 1)) {
            hipLaunchKernelGGL(k_kplane_set, dim3(blocks_for(np, TPB)), dim3(TPB), 0, st, g, 1, c->p.flux_in_sum, scal + 6,
                               c->f[VFS_UCAT], c->f[VFS_ZET], c->f[VFS_UCONT]);
            changed = true;
        }
    }
    VFS_LAUNCH_CHECK("far field / characteristic outlet");
    if (changed) ucat_stale(c);
    return VFS_OK;
}

static int form_bcs_kminus_outlet(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    double *part = c->red, *scal = c->red + 3 * RED_BLOCKS + 16;
    hipLaunchKernelGGL(k_kplane_sums, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, 0, c->f[VFS_UCAT], c->f[VFS_UCONT],
                       c->f[VFS_ZET], c->f[VFS_NVERT], part);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, part, RED_BLOCKS, scal);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, part + RED_BLOCKS, RED_BLOCKS, scal + 1);
    const long long np = (long long)(g.mx - 2) * (g.my - 2);
    hipLaunchKernelGGL(k_kplane_set, dim3(blocks_for(np, TPB)), dim3(TPB), 0, st, g, 0, c->p.flux_in_sum, scal,
                       c->f[VFS_UCAT], c->f[VFS_ZET], c->f[VFS_UCONT]);
    VFS_LAUNCH_CHECK("k- outlet");
    ucat_stale(c);
    return VFS_OK;
}

bool form_bcs_reads_ucat(const vfs_ctx *c)
{
    const Grid &g = c->g;
    bool r = g.bc[5] == 8 || g.bc[4] == 4;  // characteristic / k- outlets (k_kplane_sums, kminus)
    for (int f = 0; f < 6; f++) r = r || g.bc[f] == 6;  // far field
    return r;
}

int form_bcs(vfs_ctx *c, int outflow_scale, bool lazy_ucat)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    // lazy_ucat: Ucat is left to the next Contra2Cart (the caller's next use
    // recomputes it), so the two Contra2Cart calls are skipped; only where no
    // boundary code here reads Ucat
    const bool lazy = lazy_ucat && !form_bcs_reads_ucat(c);
    int rc = lazy ? VFS_OK : c2c(c);
    if (rc) return rc;
    rc = form_bcs_open(c);  // bcs.c:1745-2041
    if (rc) return rc;
    if (g.bc[5] == 4 && !g.pk) {
        hipLaunchKernelGGL(k_outlet_sums, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, c->f[VFS_UCONT],
                           c->f[VFS_ZET], c->f[VFS_NVERT], c->red);
        hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + 2 * RED_BLOCKS);
        hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red + RED_BLOCKS, RED_BLOCKS,
                           c->red + 2 * RED_BLOCKS + 1);
        VFS_LAUNCH_CHECK("outlet sums");
        if (outflow_scale) {
            long long n = (long long)(g.mx - 2) * (g.my - 2);
            hipLaunchKernelGGL(k_outlet_scale, dim3(blocks_for(n, TPB)), dim3(TPB), 0, st, g, c->p.flux_in_sum,
                               c->red + 2 * RED_BLOCKS, c->f[VFS_ZET], c->f[VFS_NVERT], c->f[VFS_UCONT]);
            ucat_stale(c);
            VFS_LAUNCH_CHECK("outlet scale");
        }
    }
    if (g.bc[4] == 4) {  // bcs.c:2294-2340
        rc = form_bcs_kminus_outlet(c);
        if (rc) return rc;
    }
    long long nw = (long long)(g.my > g.mx ? g.my : g.mx) * g.mz;
    hipLaunchKernelGGL(k_wall_flux, dim3(blocks_for(nw, TPB)), dim3(TPB), 0, st, g, c->f[VFS_UCONT]);
    ucat_walls_written(c);
    VFS_LAUNCH_CHECK("wall flux");
    rc = lazy ? VFS_OK : c2c(c);
    if (rc) return rc;
    int ne = g.mx > g.my ? g.mx : g.my;
    if (g.mz > ne) ne = g.mz;
    hipLaunchKernelGGL(k_zero_edges, dim3(blocks_for(ne, TPB)), dim3(TPB), 0, st, g, c->f[VFS_UCAT]);
    ucat_ghosts_written(c);
    VFS_LAUNCH_CHECK("zero edges");
    return VFS_OK;
}

// ---------------------------------------------------------------------------
// a4: Compute_eddy_viscosity_LES (les.c:1143-1361)
// ---------------------------------------------------------------------------

// Smagorinsky nu_t from the velocity gradient dd[c][d] (d/dcsi, d/deta,
// d/dzet of component c) and the cell metrics, les.c:1143-1361
__device__ inline double smagorinsky(double cs2, const double dd[3][3], const double *csi, const double *eta,
                                     const double *zet, double ajc)
{
    const double cs[3] = {csi[0], csi[1], csi[2]};
    const double et[3] = {eta[0], eta[1], eta[2]};
    const double ze[3] = {zet[0], zet[1], zet[2]};
    double gr[3][3];
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) gr[c][r] = (dd[c][0] * cs[r] + dd[c][1] * et[r] + dd[c][2] * ze[r]) * ajc;
    double Sxx = 0.5 * (gr[0][0] + gr[0][0]), Sxy = 0.5 * (gr[0][1] + gr[1][0]), Sxz = 0.5 * (gr[0][2] + gr[2][0]);
    double Syx = Sxy, Syy = 0.5 * (gr[1][1] + gr[1][1]), Syz = 0.5 * (gr[1][2] + gr[2][1]);
    double Szx = Sxz, Szy = Syz, Szz = 0.5 * (gr[2][2] + gr[2][2]);
    double Sabs = sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz + Szx * Szx +
                              Szy * Szy + Szz * Szz));
    double filter = pow(1. / ajc, 1. / 3.);
    return cs2 * pow(filter, 2.0) * Sabs;
}

__global__ void k_eddy(Grid g, double cs2, const double *__restrict__ csf, const double *__restrict__ ucat,
                       const double *__restrict__ csi, const double *__restrict__ eta, const double *__restrict__ zet,
                       const double *__restrict__ aj, const double *__restrict__ nvert, double *__restrict__ nut)
{
    Decode d;
    if (!decode(g, xcd_block1() * (long long)blockDim.x + threadIdx.x, d)) return;
    const int i = d.i, j = d.j, k = d.k;
    const long long q = gidx(g, k, j, i);
    if (i < 1 || i >= g.mx - 1 || j < 1 || j >= g.my - 1 || k < 1 || k >= g.mz - 1 || nvert[q] > 1.1) {
        nut[q] = 0;
        return;
    }
    const double SOL = 0.1;
#define U(c, kk, jj, ii) ucat[3 * gidx(g, (kk), (jj), (ii)) + (c)]
#define NV(kk, jj, ii) nvert[gidx(g, (kk), (jj), (ii))]
    double dd[3][3];
    for (int c = 0; c < 3; c++) {
        if (NV(k, j, i + 1) > SOL || (!g.pi && i == g.mx - 2)) dd[c][0] = (U(c, k, j, i) - U(c, k, j, i - 1));
        else if (NV(k, j, i - 1) > SOL || (!g.pi && i == 1)) dd[c][0] = (U(c, k, j, i + 1) - U(c, k, j, i));
        else dd[c][0] = (U(c, k, j, i + 1) - U(c, k, j, i - 1)) * 0.5;
        if (NV(k, j + 1, i) > SOL || (!g.pj && j == g.my - 2)) dd[c][1] = (U(c, k, j, i) - U(c, k, j - 1, i));
        else if (NV(k, j - 1, i) > SOL || (!g.pj && j == 1)) dd[c][1] = (U(c, k, j + 1, i) - U(c, k, j, i));
        else dd[c][1] = (U(c, k, j + 1, i) - U(c, k, j - 1, i)) * 0.5;
        // k-omega.c:390-395 tests k==0 for the lower one-sided case
        if (NV(k + 1, j, i) > SOL || (!g.pk && k == g.mz - 2)) dd[c][2] = (U(c, k, j, i) - U(c, k - 1, j, i));
        else if (NV(k - 1, j, i) > SOL || (!g.pk && k == 0)) dd[c][2] = (U(c, k + 1, j, i) - U(c, k, j, i));
        else dd[c][2] = (U(c, k + 1, j, i) - U(c, k - 1, j, i)) * 0.5;
    }
#undef U
#undef NV
    nut[q] = smagorinsky(csf ? csf[q] : cs2, dd, csi + 3 * q, eta + 3 * q, zet + 3 * q, aj[q]);
}

// The same per interior point, with one thread per (i, j) of a 64 x 4 tile
// marching KM_K planes: the thread's own ucat and nvert of the planes k-1,
// k, k+1 stay in registers, so every value of those arrays leaves HBM once
// (the thread-per-point form re-read the k +- 1 planes).  The ghost cells
// get nu_t = 0 from k_zero_ghosts.
// NB: no point is blanked (every nvert 0): the blanking tests fold away;
// CSF: lCs is the per-point array csf (dynamic model), else the scalar cs2
template <bool NB, bool CSF>
__global__ void __launch_bounds__(KM_I * KM_J) k_eddy_km(Grid g, double cs2, const double *__restrict__ csf,
                                                        const double *__restrict__ ucat,
                                                        const double *__restrict__ csi, const double *__restrict__ eta,
                                                        const double *__restrict__ zet, const double *__restrict__ aj,
                                                        const double *__restrict__ nvert, double *__restrict__ nut)
{
    int i, j, k0, k1;
    if (!km_tile(g, KM_K, i, j, k0, k1)) return;
    const double SOL = 0.1;
    const long long P = g.plane;
    long long q = gidx(g, k0, j, i);
    double um[3], u0[3], up[3];
    for (int c = 0; c < 3; c++) { um[c] = ucat[3 * (q - P) + c]; u0[c] = ucat[3 * q + c]; }
#define NVR(x) (NB ? 0.0 : nvert[x])
    double nvm = NVR(q - P), nv0 = NVR(q);
    for (int k = k0; k < k1; k++, q += P) {
        for (int c = 0; c < 3; c++) up[c] = ucat[3 * (q + P) + c];
        const double nvp = NVR(q + P);
        if (nv0 > 1.1) {
            nut[q] = 0;
        } else {
            const double nve = NVR(q + 1), nvw = NVR(q - 1), nvn = NVR(q + g.mx), nvs = NVR(q - g.mx);
            double dd[3][3];
            for (int c = 0; c < 3; c++) {
                const double ue = ucat[3 * (q + 1) + c], uw = ucat[3 * (q - 1) + c];
                const double un = ucat[3 * (q + g.mx) + c], us = ucat[3 * (q - g.mx) + c];
                if (nve > SOL || (!g.pi && i == g.mx - 2)) dd[c][0] = (u0[c] - uw);
                else if (nvw > SOL || (!g.pi && i == 1)) dd[c][0] = (ue - u0[c]);
                else dd[c][0] = (ue - uw) * 0.5;
                if (nvn > SOL || (!g.pj && j == g.my - 2)) dd[c][1] = (u0[c] - us);
                else if (nvs > SOL || (!g.pj && j == 1)) dd[c][1] = (un - u0[c]);
                else dd[c][1] = (un - us) * 0.5;
                // k-omega.c:390-395 tests k==0 for the lower one-sided case
                if (nvp > SOL || (!g.pk && k == g.mz - 2)) dd[c][2] = (u0[c] - um[c]);
                else if (nvm > SOL || (!g.pk && k == 0)) dd[c][2] = (up[c] - u0[c]);
                else dd[c][2] = (up[c] - um[c]) * 0.5;
            }
            nut[q] = smagorinsky(CSF ? csf[q] : cs2, dd, csi + 3 * q, eta + 3 * q, zet + 3 * q, aj[q]);
        }
        for (int c = 0; c < 3; c++) { um[c] = u0[c]; u0[c] = up[c]; }
        nvm = nv0;
        nv0 = nvp;
    }
#undef NVR
}

// f = 0 on the six boundary planes of the DA
__global__ void k_zero_ghosts(Grid g, double *__restrict__ f)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ni = (long long)g.my * g.mz, nj = (long long)g.mx * g.mz, nk = (long long)g.mx * g.my;
    int i, j, k;
    if (t < 2 * ni) {
        j = (int)((t >> 1) % g.my); k = (int)((t >> 1) / g.my); i = (t & 1) ? g.mx - 1 : 0;
    } else if ((t -= 2 * ni) < 2 * nj) {
        i = (int)((t >> 1) % g.mx); k = (int)((t >> 1) / g.mx); j = (t & 1) ? g.my - 1 : 0;
    } else if ((t -= 2 * nj) < 2 * nk) {
        i = (int)((t >> 1) % g.mx); j = (int)((t >> 1) / g.mx); k = (t & 1) ? g.mz - 1 : 0;
    } else {
        return;
    }
    f[((long long)k * g.my + j) * g.mx + i] = 0;
}

int zero_ghosts(vfs_ctx *c, double *f)
{
    const Grid &g = c->g;
    const long long n = 2LL * ((long long)g.my * g.mz + (long long)g.mx * g.mz + (long long)g.mx * g.my);
    hipLaunchKernelGGL(k_zero_ghosts, dim3(blocks_for(n, TPB)), dim3(TPB), 0, c->stream, g, f);
    VFS_LAUNCH_CHECK("zero ghosts");
    return VFS_OK;
}

int eddy_viscosity(vfs_ctx *c)
{
    const Grid &g = c->g;
    if (!c->p.les) {
        VFS_HIP(hipMemsetAsync(c->f[VFS_NU_T], 0, (size_t)g.n * sizeof(double), c->stream));
        return VFS_OK;
    }
    // lCs: the per-point array of the dynamic model, else one value
    const bool csf = c->cs_mode == CS_FIELD;
    const double cs2 = c->cs_mode == CS_ZERO ? 0.0 : c->p.cs2;
    const double *csp = csf ? c->f[VFS_CS] : nullptr;
    prof_begin(c, K_EDDY);
    const char *e = std::getenv("VFS_EDDY_KM");
    if (e && e[0] == '0') {
        hipLaunchKernelGGL(k_eddy, dim3(blocks_for(g.n, TPB)), dim3(TPB), 0, c->stream, g, cs2, csp, c->f[VFS_UCAT],
                           c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_NVERT], c->f[VFS_NU_T]);
    } else {
        int rc = zero_ghosts(c, c->f[VFS_NU_T]);
        if (rc) return rc;
        bool nb = false;
        rc = nvert_all_zero(c, &nb);
        if (rc) return rc;
        auto kern = nb ? (csf ? k_eddy_km<true, true> : k_eddy_km<true, false>)
                       : (csf ? k_eddy_km<false, true> : k_eddy_km<false, false>);
        hipLaunchKernelGGL(kern, dim3(km_blocks(g, KM_K)), dim3(KM_I * KM_J), 0, c->stream, g, cs2, csp,
                           c->f[VFS_UCAT], c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_NVERT],
                           c->f[VFS_NU_T]);
    }
    prof_end(c);
    VFS_LAUNCH_CHECK("eddy viscosity");
    return periodic_fix(c, c->f[VFS_NU_T], 1, 0);
}

// ---------------------------------------------------------------------------
// Calc_k_Flux (rhs2.c:1549-1633), mean_k_area (bcs.c:137-199), min dt
// ---------------------------------------------------------------------------

// mode 0: sum of ucont.z over interior k faces; 1: sum of |zet|; 2: min of
// |1/(ucont.c * aj)| over interior cells (Calc_Minimum_dt, momentum.c:157-161)
__global__ void k_interior_reduce(Grid g, int mode, const double *__restrict__ ucont, const double *__restrict__ zet,
                                  const double *__restrict__ aj, const double *__restrict__ nvert,
                                  double *__restrict__ part)
{
    const long long ni = (long long)(g.mx - 2) * (g.my - 2) * (g.mz - 2);
    double v = mode == 2 ? 1.e7 : 0.0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ni; t += (long long)gridDim.x * blockDim.x) {
        int i = 1 + t % (g.mx - 2);
        long long r = t / (g.mx - 2);
        int j = 1 + r % (g.my - 2), k = 1 + r / (g.my - 2);
        long long q = gidx(g, k, j, i);
        if (mode == 2) {
            double a = aj[q];
            for (int cc = 0; cc < 3; cc++) v = fmin(fabs(1. / ucont[3 * q + cc] / a), v);
        } else if (nvert[gidx(g, k - 1, j, i)] + nvert[q] < 0.1) {
            if (mode == 0) v += ucont[3 * q + 2];
            else v += sqrt(zet[3 * q] * zet[3 * q] + zet[3 * q + 1] * zet[3 * q + 1] + zet[3 * q + 2] * zet[3 * q + 2]);
        }
    }
    v = mode == 2 ? block_reduce<OpMin>(v) : block_reduce<OpSum>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

static int interior_reduce(vfs_ctx *c, int mode, double *out)
{
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_interior_reduce, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, c->g, mode, c->f[VFS_UCONT],
                       c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_NVERT], c->red);
    if (mode == 2) hipLaunchKernelGGL(finish_reduce<OpMin>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + RED_BLOCKS);
    else hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + RED_BLOCKS);
    VFS_LAUNCH_CHECK("interior reduce");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    *out = c->red_host[0];
    return VFS_OK;
}

int calc_k_flux(vfs_ctx *c, double *out)
{
    double s;
    int rc = interior_reduce(c, 0, &s);
    if (rc) return rc;
    *out = s / (c->g.mz - 2);
    return VFS_OK;
}

int mean_k_area(vfs_ctx *c, double *out)
{
    double s;
    int rc = interior_reduce(c, 1, &s);
    if (rc) return rc;
    *out = s / (c->g.mz - 2);
    return VFS_OK;
}

int min_dt(vfs_ctx *c, double *out) { return interior_reduce(c, 2, out); }

// ---------------------------------------------------------------------------
// a5: Pressure_Gradient (momentum.c:203-439)
// ---------------------------------------------------------------------------

// NB: no point is blanked (every nvert 0): the blanking tests fold away
template <bool NB>
__global__ void k_pgrad(Grid g, double fdiff, double fdt, double farea, const double *__restrict__ p, const double *__restrict__ nvert,
                        const double *__restrict__ icsi, const double *__restrict__ ieta, const double *__restrict__ izet,
                        const double *__restrict__ iaj, const double *__restrict__ jcsi, const double *__restrict__ jeta,
                        const double *__restrict__ jzet, const double *__restrict__ jaj, const double *__restrict__ kcsi,
                        const double *__restrict__ keta, const double *__restrict__ kzet, const double *__restrict__ kaj,
                        double *__restrict__ dp)
{
    Decode d;
    if (!decode(g, xcd_block1() * (long long)blockDim.x + threadIdx.x, d)) return;
    const int i = d.i, j = d.j, k = d.k;
    const long long q = gidx(g, k, j, i);
    const int mx = g.mx, my = g.my, mz = g.mz;
    if (i < 1 || i >= mx - 1 || j < 1 || j >= my - 1 || k < 1 || k >= mz - 1) {
        dp[3 * q] = dp[3 * q + 1] = dp[3 * q + 2] = 0;
        return;
    }
#define P(kk, jj, ii) p[gidx(g, (kk), (jj), (ii))]
#define N1(kk, jj, ii) (!NB && (int)(nvert[gidx(g, (kk), (jj), (ii))] + 0.5) == 1)
#define NVS(kk, jj, ii) (NB ? 0.0 : nvert[gidx(g, (kk), (jj), (ii))])
    const int pi_ = g.pi, pj = g.pj, pk = g.pk;
    // the three face-metric families of the point
    double FI[10], FJ[10], FK[10];
    for (int c = 0; c < 3; c++) {
        FI[c] = icsi[3 * q + c]; FI[3 + c] = ieta[3 * q + c]; FI[6 + c] = izet[3 * q + c];
        FJ[c] = jcsi[3 * q + c]; FJ[3 + c] = jeta[3 * q + c]; FJ[6 + c] = jzet[3 * q + c];
        FK[c] = kcsi[3 * q + c]; FK[3 + c] = keta[3 * q + c]; FK[6 + c] = kzet[3 * q + c];
    }
    FI[9] = iaj[q]; FJ[9] = jaj[q]; FK[9] = kaj[q];
    const double *ic = FI, *ie = FI + 3, *iz = FI + 6;
    const double *jc = FJ, *je = FJ + 3, *jz = FJ + 6;
    const double *kc = FK, *ke = FK + 3, *kz = FK + 6;
    double g11_i = ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2];
    double g12_i = ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2];
    double g13_i = iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2];
    double g21_j = jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2];
    double g22_j = je[0] * je[0] + je[1] * je[1] + je[2] * je[2];
    double g23_j = jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2];
    double g31_k = kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2];
    double g32_k = ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2];
    double g33_k = kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2];
    double dpdc, dpde, dpdz;
    if (i == mx - 2 && pi_) dpdc = P(k, j, mx + 1) - P(k, j, i);
    else dpdc = P(k, j, i + 1) - P(k, j, i);
    if (N1(k, j + 1, i) || N1(k, j + 1, i + 1) || (j == my - 2 && !pj))
        dpde = (P(k, j, i) - P(k, j - 1, i) + P(k, j, i + 1) - P(k, j - 1, i + 1)) * 0.5;
    else if (N1(k, j - 1, i) || N1(k, j - 1, i + 1) || (j == 1 && !pj))
        dpde = (P(k, j + 1, i) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j, i + 1)) * 0.5;
    else
        dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k, j + 1, i + 1) - P(k, j - 1, i + 1)) * 0.25;
    if (N1(k + 1, j, i) || N1(k + 1, j, i + 1) || (k == mz - 2 && !pk))
        dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j, i + 1) - P(k - 1, j, i + 1)) * 0.5;
    else if (N1(k - 1, j, i) || N1(k - 1, j, i + 1) || (k == 1 && !pk))
        dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j, i + 1) - P(k, j, i + 1)) * 0.5;
    else
        dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j, i + 1) - P(k - 1, j, i + 1)) * 0.25;
    double vx = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * FI[9];
    if (N1(k, j, i + 1) || N1(k, j + 1, i + 1) || (i == mx - 2 && !pi_))
        dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k, j + 1, i) - P(k, j + 1, i - 1)) * 0.5;
    else if (N1(k, j, i - 1) || N1(k, j + 1, i - 1) || (i == 1 && !pi_))
        dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j + 1, i)) * 0.5;
    else
        dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k, j + 1, i + 1) - P(k, j + 1, i - 1)) * 0.25;
    if (j == my - 2 && pj) dpde = P(k, my + 1, i) - P(k, j, i);
    else dpde = P(k, j + 1, i) - P(k, j, i);
    if (N1(k + 1, j, i) || N1(k + 1, j + 1, i) || (k == mz - 2 && !pk))
        dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j + 1, i) - P(k - 1, j + 1, i)) * 0.5;
    else if (N1(k - 1, j, i) || N1(k - 1, j + 1, i) || (k == 1 && !pk))
        dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k, j + 1, i)) * 0.5;
    else
        dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j + 1, i) - P(k - 1, j + 1, i)) * 0.25;
    double vy = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * FJ[9];
    if (N1(k, j, i + 1) || N1(k + 1, j, i + 1) || (i == mx - 2 && !pi_))
        dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k + 1, j, i) - P(k + 1, j, i - 1)) * 0.5;
    else if (N1(k, j, i - 1) || N1(k + 1, j, i - 1) || (i == 1 && !pi_))
        dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k + 1, j, i + 1) - P(k + 1, j, i)) * 0.5;
    else
        dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k + 1, j, i + 1) - P(k + 1, j, i - 1)) * 0.25;
    if (N1(k, j + 1, i) || N1(k + 1, j + 1, i) || (j == my - 2 && !pj))
        dpde = (P(k, j, i) - P(k, j - 1, i) + P(k + 1, j, i) - P(k + 1, j - 1, i)) * 0.5;
    else if (N1(k, j - 1, i) || N1(k + 1, j - 1, i) || (j == 1 && !pj))
        dpde = (P(k, j + 1, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k + 1, j, i)) * 0.5;
    else
        dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k + 1, j + 1, i) - P(k + 1, j - 1, i)) * 0.25;
    if (k == mz - 2 && pk) dpdz = P(mz + 1, j, i) - P(k, j, i);
    else dpdz = (P(k + 1, j, i) - P(k, j, i));
    if (pk) {
        double dz = 1. / FK[9] / sqrt(kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
        dpdz += dz * fdiff / fdt / farea;
    }
    double vz = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * FK[9];
    if (i == 0 || NVS(k, j, i) + NVS(k, j, i + 1) > 0.1 || (!pi_ && i == mx - 2)) vx = 0;
    if (j == 0 || NVS(k, j, i) + NVS(k, j + 1, i) > 0.1 || (!pj && j == my - 2)) vy = 0;
    if (k == 0 || NVS(k, j, i) + NVS(k + 1, j, i) > 0.1 || (!pk && k == mz - 2)) vz = 0;
    dp[3 * q] = vx; dp[3 * q + 1] = vy; dp[3 * q + 2] = vz;
#undef P
#undef N1
#undef NVS
}

int pressure_gradient(vfs_ctx *c)
{
    const Grid &g = c->g;
    int rc = periodic_fix(c, c->f[VFS_P], 1, 0);  // momentum.c:249-279
    if (rc) return rc;
    // (mean_k_flux - inlet_flux) / dt / mean_k_area, momentum.c:386-399
    const double fdiff = c->mean_k_flux - c->p.inlet_flux;
    bool nb = false;
    rc = nvert_all_zero(c, &nb);
    if (rc) return rc;
    prof_begin(c, K_PGRAD);
    {
        hipLaunchKernelGGL(nb ? k_pgrad<true> : k_pgrad<false>, dim3(blocks_for(g.n, TPB)), dim3(TPB), 0, c->stream, g, fdiff,
                           c->p.dt, c->mean_k_area,
                           c->f[VFS_P], c->f[VFS_NVERT], c->f[VFS_ICSI], c->f[VFS_IETA], c->f[VFS_IZET], c->f[VFS_IAJ],
                           c->f[VFS_JCSI], c->f[VFS_JETA], c->f[VFS_JZET], c->f[VFS_JAJ], c->f[VFS_KCSI], c->f[VFS_KETA],
                           c->f[VFS_KZET], c->f[VFS_KAJ], c->f[VFS_DP]);
    }
    prof_end(c);
    VFS_LAUNCH_CHECK("pressure gradient");
    return VFS_OK;
}

}  // namespace vfs
