// projection.hip -- Poisson right-hand side, pressure update, projection and
// divergence check (poisson_hypre.c:118-407, poisson.c:98-186, 2700-3297,
// main.c:936-1074) as HIP kernels for gfx950.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "vfs_internal.h"
#include "vfs_reduce.h"

#pragma clang fp contract(off)

namespace vfs {

constexpr int QTPB = 256;
int periodic_fix_grid(hipStream_t st, const Grid &g, double *f, int nc, int mode);

// ---------------------------------------------------------------------------
// a11: PoissonRHS2_hypre (poisson_hypre.c:265-407)
// ---------------------------------------------------------------------------

__global__ void k_poisson_rhs(Grid g, double fac, const double *__restrict__ ucont, const double *__restrict__ nvert,
                              const double *__restrict__ aj, double *__restrict__ b, double *__restrict__ part)
{
    double sb = 0, sa = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < g.n; t += (long long)gridDim.x * blockDim.x) {
        const int i = t % g.mx;
        const long long r = t / g.mx;
        const int j = r % g.my, k = r / g.my;
        double val = 0;
        if (i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1 && nvert[t] < 0.1) {
            val -= ucont[3 * t + 0];
            if (i == 1 && g.pi) val += ucont[3 * gidx(g, k, j, -2) + 0];
            else val += ucont[3 * gidx(g, k, j, i - 1) + 0];
            val -= ucont[3 * t + 1];
            if (j == 1 && g.pj) val += ucont[3 * gidx(g, k, -2, i) + 1];
            else val += ucont[3 * gidx(g, k, j - 1, i) + 1];
            val -= ucont[3 * t + 2];
            if (k == 1 && g.pk) val += ucont[3 * gidx(g, -2, j, i) + 2];
            else val += ucont[3 * gidx(g, k - 1, j, i) + 2];
            val *= fac;
            sb += val;
            sa += aj[t];
        }
        b[t] = val;
    }
    sb = block_reduce<OpSum>(sb);
    sa = block_reduce<OpSum>(sa);
    if (threadIdx.x == 0) { part[blockIdx.x] = sb; part[RED_BLOCKS + blockIdx.x] = sa; }
}

// The same over the interior with k-marching tiles (ucont.z of plane k-1
// carried in registers); the ghost cells of b get 0 from k_zero_ghosts.
__global__ void __launch_bounds__(KM_I * KM_J) k_poisson_rhs_km(Grid g, int kc, double fac, const double *__restrict__ ucont,
                                                               const double *__restrict__ nvert,
                                                               const double *__restrict__ aj, double *__restrict__ b,
                                                               double *__restrict__ part)
{
    double sb = 0, sa = 0;
    int i, j, k0, k1;
    if (km_tile(g, kc, i, j, k0, k1)) {
        const long long P = g.plane;
        long long t = gidx(g, k0, j, i);
        double wz = k0 == 1 && g.pk ? ucont[3 * gidx(g, -2, j, i) + 2] : ucont[3 * (t - P) + 2];
        for (int k = k0; k < k1; k++, t += P) {
            const double uz = ucont[3 * t + 2];
            double val = 0;
            if (nvert[t] < 0.1) {
                val -= ucont[3 * t + 0];
                if (i == 1 && g.pi) val += ucont[3 * gidx(g, k, j, -2) + 0];
                else val += ucont[3 * (t - 1) + 0];
                val -= ucont[3 * t + 1];
                if (j == 1 && g.pj) val += ucont[3 * gidx(g, k, -2, i) + 1];
                else val += ucont[3 * (t - g.mx) + 1];
                val -= uz;
                val += wz;
                val *= fac;
                sb += val;
                sa += aj[t];
            }
            b[t] = val;
            wz = uz;
        }
    }
    sb = block_reduce<OpSum>(sb);
    sa = block_reduce<OpSum>(sa);
    if (threadIdx.x == 0) { part[blockIdx.x] = sb; part[RED_BLOCKS + blockIdx.x] = sa; }
}

// planes per block of a k-marching reduction so that the blocks fit the
// RED_BLOCKS partials (0 when even whole columns do not)
static int km_chunk(const Grid &g)
{
    const long long cols = (long long)((g.mx - 2 + KM_I - 1) / KM_I) * ((g.my - 2 + KM_J - 1) / KM_J);
    if (cols > RED_BLOCKS) return 0;
    int kc = 1;
    while (cols * ((g.mz - 2 + kc - 1) / kc) > RED_BLOCKS) kc++;
    return kc;
}

static bool km_reductions()
{
    const char *e = std::getenv("VFS_REDUCE_KM");
    return !(e && e[0] == '0');
}

// Remove_Nullspace_Scale, poisson_hypre.c:118-200: b -= sum(b) * aj / sum(aj)
__global__ void k_null_scale(Grid g, const double *__restrict__ sums, const double *__restrict__ nvert,
                             const double *__restrict__ aj, double *__restrict__ b)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int i = t % g.mx;
    const long long r = t / g.mx;
    const int j = r % g.my, k = r / g.my;
    if (i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1 && nvert[t] < 0.1)
        b[t] += -sums[0] * aj[t] / sums[1];
}

int poisson_rhs(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    const double coeff = 1.;  // time_coeff(), poisson.c:23-31
    const double fac = -1.0 / c->p.dt * c->p.st * coeff;
    const int kc = km_reductions() ? km_chunk(g) : 0;
    int nparts = RED_BLOCKS;
    if (kc > 0) {
        int rc = zero_ghosts(c, c->f[VFS_B]);
        if (rc) return rc;
        nparts = (int)km_blocks(g, kc);
        hipLaunchKernelGGL(k_poisson_rhs_km, dim3(nparts), dim3(KM_I * KM_J), 0, st, g, kc, fac, c->f[VFS_UCONT],
                           c->f[VFS_NVERT], c->f[VFS_AJ], c->f[VFS_B], c->red);
    } else {
        hipLaunchKernelGGL(k_poisson_rhs, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, fac, c->f[VFS_UCONT],
                           c->f[VFS_NVERT], c->f[VFS_AJ], c->f[VFS_B], c->red);
    }
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, nparts, c->red + 2 * RED_BLOCKS);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red + RED_BLOCKS, nparts,
                       c->red + 2 * RED_BLOCKS + 1);
    hipLaunchKernelGGL(k_null_scale, dim3(blocks_for(g.n, QTPB)), dim3(QTPB), 0, st, g, c->red + 2 * RED_BLOCKS,
                       c->f[VFS_NVERT], c->f[VFS_AJ], c->f[VFS_B]);
    VFS_LAUNCH_CHECK("poisson rhs");
    return VFS_OK;
}

// ---------------------------------------------------------------------------
// a13: Convert_Phi2_Phi + UpdatePressure (poisson.c:98-186, 3137-3297)
// ---------------------------------------------------------------------------

__global__ void k_phi_mask(Grid g, double th, const double *__restrict__ nvert, double *__restrict__ phi)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int i = t % g.mx;
    const long long r = t / g.mx;
    const int j = r % g.my, k = r / g.my;
    bool rowc = i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1 && nvert[t] < th;
    if (!rowc) phi[t] = 0;
}

__global__ void k_update_p(Grid g, const double *__restrict__ nvert, const double *__restrict__ phi, double *__restrict__ p)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int i = t % g.mx;
    const long long r = t / g.mx;
    const int j = r % g.my, k = r / g.my;
    if (i < 1 || i >= g.mx - 1 || j < 1 || j >= g.my - 1 || k < 1 || k >= g.mz - 1) return;
    if (nvert[t] < 0.1) p[t] += phi[t];
    if (nvert[t] > 1.1) p[t] = 0;
}

int update_pressure(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    const unsigned nb = blocks_for(g.n, QTPB);
    hipLaunchKernelGGL(k_phi_mask, dim3(nb), dim3(QTPB), 0, st, g, c->p.poisson_threshold, c->f[VFS_NVERT], c->f[VFS_PHI]);
    VFS_LAUNCH_CHECK("phi mask");
    int rc = periodic_fix_grid(st, g, c->f[VFS_PHI], 1, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_update_p, dim3(nb), dim3(QTPB), 0, st, g, c->f[VFS_NVERT], c->f[VFS_PHI], c->f[VFS_P]);
    VFS_LAUNCH_CHECK("update pressure");
    return periodic_fix_grid(st, g, c->f[VFS_P], 1, 0);
}

// ---------------------------------------------------------------------------
// a14: Projection (poisson.c:2700-3066)
// ---------------------------------------------------------------------------

struct ProjArgs {
    Grid g;
    double th, fac;  // poisson_threshold; dt*st/coeff
    const double *phi, *nvert;
    const double *icsi, *ieta, *izet, *iaj, *jcsi, *jeta, *jzet, *jaj, *kcsi, *keta, *kzet, *kaj;
    double *ucont;
};

// NB: no point is blanked (every nvert 0): NV is the constant 0
template <bool NB>
__global__ void __launch_bounds__(QTPB) k_projection(ProjArgs a)
{
    const Grid &g = a.g;
    unsigned bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by + 1, k = bz + 1;
    if (i >= g.mx - 1) return;
    const int mx = g.mx, my = g.my, mz = g.mz, pi_ = g.pi, pj = g.pj, pk = g.pk;
    const double th = a.th;
    const long long q = gidx(g, k, j, i);
#define PH(kk, jj, ii) a.phi[gidx(g, (kk), (jj), (ii))]
#define NV(kk, jj, ii) (NB ? 0.0 : a.nvert[gidx(g, (kk), (jj), (ii))])
    // the three components are read and written together (separate
    // read-modify-writes of one component of the AoS line each evicted
    // partly written lines)
    double u[3] = {a.ucont[3 * q + 0], a.ucont[3 * q + 1], a.ucont[3 * q + 2]};
    double dpdc, dpde, dpdz;
    if (i < mx - 2 || (pi_ && i == mx - 2)) {
        dpdc = PH(k, j, i + 1) - PH(k, j, i);
        if (i == mx - 2 && pi_) dpdc = PH(k, j, mx + 1) - PH(k, j, i);
        dpde = 0.; dpdz = 0.;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i + 1) < th && (j != 1))
                dpde = (PH(k, j, i) + PH(k, j, i + 1) - PH(k, j - 1, i) - PH(k, j - 1, i + 1)) * 0.5;
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i + 1) < th)
                dpde = (PH(k, j + 1, i) + PH(k, j + 1, i + 1) - PH(k, j, i) - PH(k, j, i + 1)) * 0.5;
        } else
            dpde = (PH(k, j + 1, i) + PH(k, j + 1, i + 1) - PH(k, j - 1, i) - PH(k, j - 1, i + 1)) * 0.25;
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j, i + 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i + 1) < th && (k != 1))
                dpdz = (PH(k, j, i) + PH(k, j, i + 1) - PH(k - 1, j, i) - PH(k - 1, j, i + 1)) * 0.5;
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j, i + 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i + 1) < th)
                dpdz = (PH(k + 1, j, i) + PH(k + 1, j, i + 1) - PH(k, j, i) - PH(k, j, i + 1)) * 0.5;
        } else
            dpdz = (PH(k + 1, j, i) + PH(k + 1, j, i + 1) - PH(k - 1, j, i) - PH(k - 1, j, i + 1)) * 0.25;
        if ((NV(k, j, i) + NV(k, j, i + 1)) < th) {
            const double *ic = a.icsi + 3 * q, *ie = a.ieta + 3 * q, *iz = a.izet + 3 * q;
            const double iaj = a.iaj[q];
            u[0] -= (dpdc * (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]) * iaj +
                     dpde * (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]) * iaj +
                     dpdz * (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]) * iaj) * a.fac;
        }
    }
    if (j < my - 2 || (pj && j == my - 2)) {
        dpdc = 0.; dpdz = 0.;
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j + 1, i - 1) < th && (i != 1))
                dpdc = (PH(k, j, i) + PH(k, j + 1, i) - PH(k, j, i - 1) - PH(k, j + 1, i - 1)) * 0.5;
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j + 1, i + 1) < th)
                dpdc = (PH(k, j, i + 1) + PH(k, j + 1, i + 1) - PH(k, j, i) - PH(k, j + 1, i)) * 0.5;
        } else
            dpdc = (PH(k, j, i + 1) + PH(k, j + 1, i + 1) - PH(k, j, i - 1) - PH(k, j + 1, i - 1)) * 0.25;
        if (j == my - 2 && pj) dpde = PH(k, my + 1, i) - PH(k, j, i);
        else dpde = PH(k, j + 1, i) - PH(k, j, i);
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j + 1, i) < th && (k != 1))
                dpdz = (PH(k, j, i) + PH(k, j + 1, i) - PH(k - 1, j, i) - PH(k - 1, j + 1, i)) * 0.5;
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j + 1, i) < th)
                dpdz = (PH(k + 1, j, i) + PH(k + 1, j + 1, i) - PH(k, j, i) - PH(k, j + 1, i)) * 0.5;
        } else
            dpdz = (PH(k + 1, j, i) + PH(k + 1, j + 1, i) - PH(k - 1, j, i) - PH(k - 1, j + 1, i)) * 0.25;
        if ((NV(k, j, i) + NV(k, j + 1, i)) < th || (int)(NV(k, j, i) + NV(k, j + 1, i)) == 5) {
            const double *jc = a.jcsi + 3 * q, *je = a.jeta + 3 * q, *jz = a.jzet + 3 * q;
            const double jaj = a.jaj[q];
            u[1] -= (dpdc * (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]) * jaj +
                     dpde * (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]) * jaj +
                     dpdz * (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]) * jaj) * a.fac;
        }
    }
    if (k < mz - 2 || (pk && k == mz - 2)) {
        dpdc = 0.; dpde = 0.;
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k + 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k + 1, j, i - 1) < th && (i != 1))
                dpdc = (PH(k, j, i) + PH(k + 1, j, i) - PH(k, j, i - 1) - PH(k + 1, j, i - 1)) * 0.5;
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k + 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k + 1, j, i + 1) < th)
                dpdc = (PH(k, j, i + 1) + PH(k + 1, j, i + 1) - PH(k, j, i) - PH(k + 1, j, i)) * 0.5;
        } else
            dpdc = (PH(k, j, i + 1) + PH(k + 1, j, i + 1) - PH(k, j, i - 1) - PH(k + 1, j, i - 1)) * 0.25;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k + 1, j - 1, i) < th && (j != 1))
                dpde = (PH(k, j, i) + PH(k + 1, j, i) - PH(k, j - 1, i) - PH(k + 1, j - 1, i)) * 0.5;
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k + 1, j + 1, i) < th)
                dpde = (PH(k, j + 1, i) + PH(k + 1, j + 1, i) - PH(k, j, i) - PH(k + 1, j, i)) * 0.5;
        } else
            dpde = (PH(k, j + 1, i) + PH(k + 1, j + 1, i) - PH(k, j - 1, i) - PH(k + 1, j - 1, i)) * 0.25;
        if (k == mz - 2 && pk) dpdz = PH(mz + 1, j, i) - PH(k, j, i);
        else dpdz = PH(k + 1, j, i) - PH(k, j, i);
        if ((NV(k, j, i) + NV(k + 1, j, i)) < th) {
            const double *kc = a.kcsi + 3 * q, *ke = a.keta + 3 * q, *kz = a.kzet + 3 * q;
            const double kaj = a.kaj[q];
            u[2] -= (dpdc * (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]) * kaj +
                     dpde * (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]) * kaj +
                     dpdz * (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]) * kaj) * a.fac;
        }
    }
    a.ucont[3 * q + 0] = u[0];
    a.ucont[3 * q + 1] = u[1];
    a.ucont[3 * q + 2] = u[2];
#undef PH
#undef NV
}

int projection(vfs_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    ProjArgs a;
    a.g = g;
    a.th = c->p.poisson_threshold;
    const double coeff = 1.;
    a.fac = c->p.dt * c->p.st / coeff;
    a.phi = c->f[VFS_PHI];
    a.nvert = c->f[VFS_NVERT];
    a.icsi = c->f[VFS_ICSI]; a.ieta = c->f[VFS_IETA]; a.izet = c->f[VFS_IZET]; a.iaj = c->f[VFS_IAJ];
    a.jcsi = c->f[VFS_JCSI]; a.jeta = c->f[VFS_JETA]; a.jzet = c->f[VFS_JZET]; a.jaj = c->f[VFS_JAJ];
    a.kcsi = c->f[VFS_KCSI]; a.keta = c->f[VFS_KETA]; a.kzet = c->f[VFS_KZET]; a.kaj = c->f[VFS_KAJ];
    a.ucont = c->f[VFS_UCONT];
    dim3 grd(blocks_for(g.mx - 2, QTPB), g.my - 2, g.mz - 2);
    bool nb = false;
    int rc0 = nvert_all_zero(c, &nb);
    if (rc0) return rc0;
    if (nb) hipLaunchKernelGGL(k_projection<true>, grd, dim3(QTPB), 0, st, a);
    else hipLaunchKernelGGL(k_projection<false>, grd, dim3(QTPB), 0, st, a);
    ucat_stale(c);
    VFS_LAUNCH_CHECK("projection");
    int rc = periodic_fix_grid(st, g, c->f[VFS_UCONT], 3, 1);  // poisson.c:3001-3040
    if (rc) return rc;
    return c2c(c);  // poisson.c:3061
}

// ---------------------------------------------------------------------------
// a15: Divergence (main.c:936-1074)
// ---------------------------------------------------------------------------

__global__ void k_divergence(Grid g, const double *__restrict__ ucont, const double *__restrict__ aj,
                             const double *__restrict__ nvert, double *__restrict__ part)
{
    const long long ni = (long long)(g.mx - 2) * (g.my - 2) * (g.mz - 2);
    double v = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ni; t += (long long)gridDim.x * blockDim.x) {
        int i = 1 + t % (g.mx - 2);
        long long r = t / (g.mx - 2);
        int j = 1 + r % (g.my - 2), k = 1 + r / (g.my - 2);
        long long q = gidx(g, k, j, i);
        double d = fabs((ucont[3 * q] - ucont[3 * gidx(g, k, j, i - 1)] + ucont[3 * q + 1] - ucont[3 * gidx(g, k, j - 1, i) + 1] +
                         ucont[3 * q + 2] - ucont[3 * gidx(g, k - 1, j, i) + 2]) * aj[q]);
        if (nvert[q] + nvert[gidx(g, k + 1, j, i)] + nvert[gidx(g, k - 1, j, i)] + nvert[gidx(g, k, j + 1, i)] +
                nvert[gidx(g, k, j - 1, i)] + nvert[gidx(g, k, j, i + 1)] + nvert[gidx(g, k, j, i - 1)] > 0.1)
            d = 0.;
        v = fmax(v, d);
    }
    v = block_reduce<OpMax>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// The same with k-marching tiles: ucont.z and nvert of the planes k-1, k,
// k+1 at the thread's own (i, j) carried in registers
__global__ void __launch_bounds__(KM_I * KM_J) k_divergence_km(Grid g, int kc, const double *__restrict__ ucont,
                                                              const double *__restrict__ aj,
                                                              const double *__restrict__ nvert, double *__restrict__ part)
{
    double v = 0;
    int i, j, k0, k1;
    if (km_tile(g, kc, i, j, k0, k1)) {
        const long long P = g.plane;
        long long q = gidx(g, k0, j, i);
        double wz = ucont[3 * (q - P) + 2], nvm = nvert[q - P], nv0 = nvert[q];
        for (int k = k0; k < k1; k++, q += P) {
            const double uz = ucont[3 * q + 2], nvp = nvert[q + P];
            double d = fabs((ucont[3 * q] - ucont[3 * (q - 1)] + ucont[3 * q + 1] - ucont[3 * (q - g.mx) + 1] + uz - wz) * aj[q]);
            if (nv0 + nvp + nvm + nvert[q + g.mx] + nvert[q - g.mx] + nvert[q + 1] + nvert[q - 1] > 0.1) d = 0.;
            v = fmax(v, d);
            wz = uz;
            nvm = nv0;
            nv0 = nvp;
        }
    }
    v = block_reduce<OpMax>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

int divergence(vfs_ctx *c, double *out)
{
    hipStream_t st = c->stream;
    const int kc = km_reductions() ? km_chunk(c->g) : 0;
    int nparts = RED_BLOCKS;
    if (kc > 0) {
        nparts = (int)km_blocks(c->g, kc);
        hipLaunchKernelGGL(k_divergence_km, dim3(nparts), dim3(KM_I * KM_J), 0, st, c->g, kc, c->f[VFS_UCONT], c->f[VFS_AJ],
                           c->f[VFS_NVERT], c->red);
    } else {
        hipLaunchKernelGGL(k_divergence, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, c->g, c->f[VFS_UCONT], c->f[VFS_AJ],
                           c->f[VFS_NVERT], c->red);
    }
    hipLaunchKernelGGL(finish_reduce<OpMax>, dim3(1), dim3(RED_THREADS), 0, st, c->red, nparts, c->red + RED_BLOCKS);
    VFS_LAUNCH_CHECK("divergence");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    *out = c->red_host[0];
    return VFS_OK;
}

// ---------------------------------------------------------------------------
// Divergence without device allocation, for callers that cannot allocate
// (code under HIP graph capture, serving loops).  The same kernels as
// divergence() write their per-block partials into `scratch`, a file-scope
// device array of NOALLOC_PARTS doubles (8 KiB) that every GPU the library
// runs on holds until the process ends.  Its device address is looked up
// once per device (hipGetSymbolAddress, on the first call with that device
// current) and only that address is handed to kernels.  Calls that overlap
// on one GPU share the array and race.
// ---------------------------------------------------------------------------

constexpr int NOALLOC_PARTS = RED_BLOCKS;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_PARTS];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static int scratch_address(double **out)
{
    int dev = -1;
    VFS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_error(VFS_EINVAL, "divergence_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return VFS_OK;
}

int divergence_noalloc(const vfs_params &p, const double *ucont, const double *aj, const double *nvert, double *maxdiv,
                       hipStream_t st)
{
    if (p.mx < 3 || p.my < 3 || p.mz < 3) return set_error(VFS_EINVAL, "divergence_noalloc: each DA size must be at least 3");
    if ((long long)p.mx * p.my * p.mz >= (1LL << 31))
        return set_error(VFS_EINVAL, "divergence_noalloc: grid too large: the DA must hold fewer than 2^31 points");
    const Grid g = make_grid(p);
    const int kc = km_reductions() ? km_chunk(g) : 0;
    const long long nparts = kc > 0 ? (long long)km_blocks(g, kc) : RED_BLOCKS;
    if (nparts > NOALLOC_PARTS)
        return set_error(VFS_EINVAL, "divergence_noalloc: " + std::to_string(nparts) + " partial results exceed the " +
                                         std::to_string(NOALLOC_PARTS) + " of the scratch array");
    double *part = nullptr;
    int rc = scratch_address(&part);
    if (rc) return rc;
    // every block writes its own partial, so nothing needs clearing between calls
    if (kc > 0)
        hipLaunchKernelGGL(k_divergence_km, dim3((unsigned)nparts), dim3(KM_I * KM_J), 0, st, g, kc, ucont, aj, nvert, part);
    else
        hipLaunchKernelGGL(k_divergence, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, g, ucont, aj, nvert, part);
    hipLaunchKernelGGL(finish_reduce<OpMax>, dim3(1), dim3(RED_THREADS), 0, st, part, (int)nparts, maxdiv);
    VFS_LAUNCH_CHECK("divergence_noalloc");
    return VFS_OK;
}

}  // namespace vfs

extern "C" int vfs_divergence_noalloc(const vfs_params *params, const double *ucont, const double *aj,
                                      const double *nvert, double *maxdiv, void *stream)
{
    if (!params || !ucont || !aj || !nvert || !maxdiv) return vfs::set_error(VFS_EINVAL, "divergence_noalloc: null argument");
    return vfs::divergence_noalloc(*params, ucont, aj, nvert, maxdiv, static_cast<hipStream_t>(stream));
}
