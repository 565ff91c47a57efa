// poisson.hip -- pressure Poisson operator, multigrid and FGMRES for gfx950.
//
// Operator: the poisson==1 rows of PoissonLHSNew (poisson.c:977-1688): one
// row per fluid interior cell, 19 coefficients assembled from the face metric
// products g_ab = (n_a . n_b) * face aj with the reference's one-sided
// cross-derivative switching at walls and blanked cells.  Coefficients are
// stored as 19 structure-of-arrays planes (coalesced), in FP64 for the outer
// residual and in FP32 for the mixed-precision preconditioner.
//
// Preconditioner: geometric V-cycle; coarse grids take every other node
// (init.c:402-455) and are rediscretised through the same metric and
// assembly kernels; restriction sums the eight child residuals (rows are flux
// sums), prolongation is cell-centred trilinear; the smoother is damped
// Jacobi over wall-normal (j) lines, each line solved exactly by the Thomas
// algorithm, one thread per line with the sweep marching in j so that every
// load coalesces across the i-consecutive lanes of a wavefront.
//
// Outer solver: FGMRES(m) in FP64, classical Gram-Schmidt with one selective
// re-orthogonalisation; the FP32 V-cycle inside it acts as iterative
// refinement, so the FP64 true residual reaches the requested tolerance.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"
#include "vfs_krylov.h"

#pragma clang fp contract(off)

namespace vfs {

constexpr int PTPB = 256;
constexpr int NST = 19;
enum { CP, EP, WP, NP, SP, TP, BP, NE, SE, NW, SW, TN, BN, TS, BS, TE, BE, TW, BW };
// stencil offsets (di, dj, dk) of each slot
__host__ __device__ constexpr int soff(int s, int d)
{
    constexpr int o[NST][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                               {0, 0, 1},  {0, 0, -1},  {1, 1, 0},   {1, -1, 0}, {-1, 1, 0},
                               {-1, -1, 0}, {0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1},
                               {1, 0, 1},  {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1}};
    return o[s][d];
}

int periodic_fix_grid(hipStream_t st, const Grid &g, double *f, int nc, int mode);

// periodic image of a neighbour index (gid ghost mapping, poisson.c:284-309)
__host__ __device__ inline int canon(int i, int m, int per)
{
    if (per) {
        if (i <= 0) i += m - 2;
        else if (i >= m - 1) i -= m - 2;
    }
    return i;
}

__host__ __device__ inline long long nbr_idx(const Grid &g, int k, int j, int i, int s)
{
    int ii = canon(i + soff(s, 0), g.mx, g.pi);
    int jj = canon(j + soff(s, 1), g.my, g.pj);
    int kk = canon(k + soff(s, 2), g.mz, g.pk);
    return ((long long)kk * g.my + jj) * g.mx + ii;
}

struct FaceSet {
    const double *icsi, *ieta, *izet, *iaj, *jcsi, *jeta, *jzet, *jaj, *kcsi, *keta, *kzet, *kaj;
};

__device__ inline void face_g(const Grid &g, const FaceSet &F, int k, int j, int i, double gg[9])
{
    long long q = gidx(g, k, j, i);
    const double *ic = F.icsi + 3 * q, *ie = F.ieta + 3 * q, *iz = F.izet + 3 * q;
    const double *jc = F.jcsi + 3 * q, *je = F.jeta + 3 * q, *jz = F.jzet + 3 * q;
    const double *kc = F.kcsi + 3 * q, *ke = F.keta + 3 * q, *kz = F.kzet + 3 * q;
    const double iaj = F.iaj[q], jaj = F.jaj[q], kaj = F.kaj[q];
    gg[0] = (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]) * iaj;
    gg[1] = (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]) * iaj;
    gg[2] = (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]) * iaj;
    gg[3] = (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]) * jaj;
    gg[4] = (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]) * jaj;
    gg[5] = (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]) * jaj;
    gg[6] = (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]) * kaj;
    gg[7] = (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]) * kaj;
    gg[8] = (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]) * kaj;
}

// the nine face metric products of point (k, j, i), precomputed planes
struct FaceG {
    const double *gp;  // 9 planes of g.n values
    long long n;
};
__device__ inline void face_g(const Grid &g, const FaceG &F, int k, int j, int i, double gg[9])
{
    const long long q = gidx(g, k, j, i);
#pragma unroll
    for (int c = 0; c < 9; c++) gg[c] = F.gp[c * F.n + q];
}

__global__ void k_face_g(Grid g, FaceSet F, double *__restrict__ gp)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int t32 = (int)t; const int i = t32 % g.mx; const int r = t32 / g.mx; const int j = r % g.my, k = r / g.my;
    double gg[9];
    face_g(g, F, k, j, i, gg);
    for (int c = 0; c < 9; c++) gp[c * g.n + t] = gg[c];
}

// one row of PoissonLHSNew (poisson.c:1159-1561), poisson==1 form; with
// NB (no point blanked) every nvert read is the constant 0
template <bool NB = false, class FS = FaceSet>
__device__ void assemble_row(const Grid &g, const FS &F, const double *nvert, double th, int k, int j, int i,
                             double vol[NST])
{
    const int mx = g.mx, my = g.my, mz = g.mz, pi_ = g.pi, pj = g.pj, pk = g.pk;
    const double r = 1.0;
    for (int s = 0; s < NST; s++) vol[s] = 0.;
#define NV(kk, jj, ii) (NB ? 0.0 : nvert[gidx(g, (kk), (jj), (ii))])
    double gg[9];
    if (NV(k, j, i + 1) < th && (i != mx - 2 || pi_)) {
        face_g(g, F, k, j, i, gg);
        vol[CP] -= gg[0] / r; vol[EP] += gg[0] / r;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i + 1) < th && j != 1) {
                vol[CP] += gg[1] * 0.5 / r; vol[EP] += gg[1] * 0.5 / r; vol[SP] -= gg[1] * 0.5 / r; vol[SE] -= gg[1] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i + 1) < th) {
                vol[NP] += gg[1] * 0.5 / r; vol[NE] += gg[1] * 0.5 / r; vol[CP] -= gg[1] * 0.5 / r; vol[EP] -= gg[1] * 0.5 / r;
            }
        } else {
            vol[NP] += gg[1] * 0.25 / r; vol[NE] += gg[1] * 0.25 / r; vol[SP] -= gg[1] * 0.25 / r; vol[SE] -= gg[1] * 0.25 / r;
        }
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j, i + 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i + 1) < th && k != 1) {
                vol[CP] += gg[2] * 0.5 / r; vol[EP] += gg[2] * 0.5 / r; vol[BP] -= gg[2] * 0.5 / r; vol[BE] -= gg[2] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j, i + 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i + 1) < th) {
                vol[TP] += gg[2] * 0.5 / r; vol[TE] += gg[2] * 0.5 / r; vol[CP] -= gg[2] * 0.5 / r; vol[EP] -= gg[2] * 0.5 / r;
            }
        } else {
            vol[TP] += gg[2] * 0.25 / r; vol[TE] += gg[2] * 0.25 / r; vol[BP] -= gg[2] * 0.25 / r; vol[BE] -= gg[2] * 0.25 / r;
        }
    }
    if (NV(k, j, i - 1) < th && (i != 1 || pi_)) {
        face_g(g, F, k, j, i - 1, gg);
        vol[CP] -= gg[0] / r; vol[WP] += gg[0] / r;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i - 1) < th && j != 1) {
                vol[CP] -= gg[1] * 0.5 / r; vol[WP] -= gg[1] * 0.5 / r; vol[SP] += gg[1] * 0.5 / r; vol[SW] += gg[1] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i - 1) < th) {
                vol[NP] -= gg[1] * 0.5 / r; vol[NW] -= gg[1] * 0.5 / r; vol[CP] += gg[1] * 0.5 / r; vol[WP] += gg[1] * 0.5 / r;
            }
        } else {
            vol[NP] -= gg[1] * 0.25 / r; vol[NW] -= gg[1] * 0.25 / r; vol[SP] += gg[1] * 0.25 / r; vol[SW] += gg[1] * 0.25 / r;
        }
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j, i - 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i - 1) < th && k != 1) {
                vol[CP] -= gg[2] * 0.5 / r; vol[WP] -= gg[2] * 0.5 / r; vol[BP] += gg[2] * 0.5 / r; vol[BW] += gg[2] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j, i - 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i - 1) < th) {
                vol[TP] -= gg[2] * 0.5 / r; vol[TW] -= gg[2] * 0.5 / r; vol[CP] += gg[2] * 0.5 / r; vol[WP] += gg[2] * 0.5 / r;
            }
        } else {
            vol[TP] -= gg[2] * 0.25 / r; vol[TW] -= gg[2] * 0.25 / r; vol[BP] += gg[2] * 0.25 / r; vol[BW] += gg[2] * 0.25 / r;
        }
    }
    if (NV(k, j + 1, i) < th && (j != my - 2 || pj)) {
        face_g(g, F, k, j, i, gg);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j + 1, i - 1) < th && i != 1) {
                vol[CP] += gg[3] * 0.5 / r; vol[NP] += gg[3] * 0.5 / r; vol[WP] -= gg[3] * 0.5 / r; vol[NW] -= gg[3] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j + 1, i + 1) < th) {
                vol[EP] += gg[3] * 0.5 / r; vol[NE] += gg[3] * 0.5 / r; vol[CP] -= gg[3] * 0.5 / r; vol[NP] -= gg[3] * 0.5 / r;
            }
        } else {
            vol[EP] += gg[3] * 0.25 / r; vol[NE] += gg[3] * 0.25 / r; vol[WP] -= gg[3] * 0.25 / r; vol[NW] -= gg[3] * 0.25 / r;
        }
        vol[CP] -= gg[4] / r; vol[NP] += gg[4] / r;
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j + 1, i) < th && k != 1) {
                vol[CP] += gg[5] * 0.5 / r; vol[NP] += gg[5] * 0.5 / r; vol[BP] -= gg[5] * 0.5 / r; vol[BN] -= gg[5] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j + 1, i) < th) {
                vol[TP] += gg[5] * 0.5 / r; vol[TN] += gg[5] * 0.5 / r; vol[CP] -= gg[5] * 0.5 / r; vol[NP] -= gg[5] * 0.5 / r;
            }
        } else {
            vol[TP] += gg[5] * 0.25 / r; vol[TN] += gg[5] * 0.25 / r; vol[BP] -= gg[5] * 0.25 / r; vol[BN] -= gg[5] * 0.25 / r;
        }
    }
    if (NV(k, j - 1, i) < th && (j != 1 || pj)) {
        face_g(g, F, k, j - 1, i, gg);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j - 1, i - 1) < th && i != 1) {
                vol[CP] -= gg[3] * 0.5 / r; vol[SP] -= gg[3] * 0.5 / r; vol[WP] += gg[3] * 0.5 / r; vol[SW] += gg[3] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j - 1, i + 1) < th) {
                vol[EP] -= gg[3] * 0.5 / r; vol[SE] -= gg[3] * 0.5 / r; vol[CP] += gg[3] * 0.5 / r; vol[SP] += gg[3] * 0.5 / r;
            }
        } else {
            vol[EP] -= gg[3] * 0.25 / r; vol[SE] -= gg[3] * 0.25 / r; vol[WP] += gg[3] * 0.25 / r; vol[SW] += gg[3] * 0.25 / r;
        }
        vol[CP] -= gg[4] / r; vol[SP] += gg[4] / r;
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j - 1, i) < th && k != 1) {
                vol[CP] -= gg[5] * 0.5 / r; vol[SP] -= gg[5] * 0.5 / r; vol[BP] += gg[5] * 0.5 / r; vol[BS] += gg[5] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j - 1, i) < th) {
                vol[TP] -= gg[5] * 0.5 / r; vol[TS] -= gg[5] * 0.5 / r; vol[CP] += gg[5] * 0.5 / r; vol[SP] += gg[5] * 0.5 / r;
            }
        } else {
            vol[TP] -= gg[5] * 0.25 / r; vol[TS] -= gg[5] * 0.25 / r; vol[BP] += gg[5] * 0.25 / r; vol[BS] += gg[5] * 0.25 / r;
        }
    }
    if (NV(k + 1, j, i) < th && (k != mz - 2 || pk)) {
        face_g(g, F, k, j, i, gg);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k + 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k + 1, j, i - 1) < th && i != 1) {
                vol[CP] += gg[6] * 0.5 / r; vol[TP] += gg[6] * 0.5 / r; vol[WP] -= gg[6] * 0.5 / r; vol[TW] -= gg[6] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k + 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k + 1, j, i + 1) < th) {
                vol[EP] += gg[6] * 0.5 / r; vol[TE] += gg[6] * 0.5 / r; vol[CP] -= gg[6] * 0.5 / r; vol[TP] -= gg[6] * 0.5 / r;
            }
        } else {
            vol[EP] += gg[6] * 0.25 / r; vol[TE] += gg[6] * 0.25 / r; vol[WP] -= gg[6] * 0.25 / r; vol[TW] -= gg[6] * 0.25 / r;
        }
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k + 1, j - 1, i) < th && j != 1) {
                vol[CP] += gg[7] * 0.5 / r; vol[TP] += gg[7] * 0.5 / r; vol[SP] -= gg[7] * 0.5 / r; vol[TS] -= gg[7] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k + 1, j + 1, i) < th) {
                vol[NP] += gg[7] * 0.5 / r; vol[TN] += gg[7] * 0.5 / r; vol[CP] -= gg[7] * 0.5 / r; vol[TP] -= gg[7] * 0.5 / r;
            }
        } else {
            vol[NP] += gg[7] * 0.25 / r; vol[TN] += gg[7] * 0.25 / r; vol[SP] -= gg[7] * 0.25 / r; vol[TS] -= gg[7] * 0.25 / r;
        }
        vol[CP] -= gg[8] / r; vol[TP] += gg[8] / r;
    }
    if (NV(k - 1, j, i) < th && (k != 1 || pk)) {
        face_g(g, F, k - 1, j, i, gg);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k - 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k - 1, j, i - 1) < th && i != 1) {
                vol[CP] -= gg[6] * 0.5 / r; vol[BP] -= gg[6] * 0.5 / r; vol[WP] += gg[6] * 0.5 / r; vol[BW] += gg[6] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k - 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k - 1, j, i + 1) < th) {
                vol[EP] -= gg[6] * 0.5 / r; vol[BE] -= gg[6] * 0.5 / r; vol[CP] += gg[6] * 0.5 / r; vol[BP] += gg[6] * 0.5 / r;
            }
        } else {
            vol[EP] -= gg[6] * 0.25 / r; vol[BE] -= gg[6] * 0.25 / r; vol[WP] += gg[6] * 0.25 / r; vol[BW] += gg[6] * 0.25 / r;
        }
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k - 1, j - 1, i) < th && j != 1) {
                vol[CP] -= gg[7] * 0.5 / r; vol[BP] -= gg[7] * 0.5 / r; vol[SP] += gg[7] * 0.5 / r; vol[BS] += gg[7] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k - 1, j + 1, i) < th) {
                vol[NP] -= gg[7] * 0.5 / r; vol[BN] -= gg[7] * 0.5 / r; vol[CP] += gg[7] * 0.5 / r; vol[BP] += gg[7] * 0.5 / r;
            }
        } else {
            vol[NP] -= gg[7] * 0.25 / r; vol[BN] -= gg[7] * 0.25 / r; vol[SP] += gg[7] * 0.25 / r; vol[BS] += gg[7] * 0.25 / r;
        }
        vol[CP] -= gg[8] / r; vol[BP] += gg[8] / r;
    }
#undef NV
}

__device__ inline bool is_row(const Grid &g, const double *nvert, double th, int k, int j, int i)
{
    if (i <= 0 || i >= g.mx - 1 || j <= 0 || j >= g.my - 1 || k <= 0 || k >= g.mz - 1) return false;
    return nvert[gidx(g, k, j, i)] < th;
}

__global__ void k_row_mask(Grid g, const double *__restrict__ nvert, double th, unsigned char *__restrict__ row)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int t32 = (int)t; int i = t32 % g.mx; int r = t32 / g.mx; int j = r % g.my, k = r / g.my;
    row[t] = is_row(g, nvert, th, k, j, i) ? 1 : 0;
}

// coefficients of every row; columns of non-rows are dropped like PETSc
// drops the negative gid columns of PoissonLHSNew's MatSetValues
__global__ void k_assemble(Grid g, FaceSet F, const double *__restrict__ nvert, double th,
                           const unsigned char *__restrict__ row, double *__restrict__ c64, float *__restrict__ c32)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    const int t32 = (int)t; int i = t32 % g.mx; int r = t32 / g.mx; int j = r % g.my, k = r / g.my;
    double vol[NST];
    if (row[t]) {
        assemble_row(g, F, nvert, th, k, j, i, vol);
        for (int s = 1; s < NST; s++)
            if (!row[nbr_idx(g, k, j, i, s)]) vol[s] = 0.;
    } else {
        for (int s = 0; s < NST; s++) vol[s] = 0.;
    }
    for (int s = 0; s < NST; s++) {
        if (c64) c64[s * g.n + t] = vol[s];
        c32[s * g.n + t] = (float)vol[s];
    }
}

// y = A x (rows only), or r = b - A x when b != nullptr; x may be stored in a
// narrower type X (an FP32 preconditioned vector, exact in FP64).  NQ = 7
// applies only the centre and the six axis neighbours (the compact levels'
// V-cycle operator; the stencil's first seven slots).
template <typename T, typename X = T, int NQ = NST>
__global__ void __launch_bounds__(PTPB) k_apply(Grid g, const T *__restrict__ coef, const unsigned char *__restrict__ row,
                                                const X *__restrict__ x, const T *__restrict__ b, T *__restrict__ y)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    if (!row[t]) { y[t] = 0; return; }
    const int t32 = (int)t; int i = t32 % g.mx; int r = t32 / g.mx; int j = r % g.my, k = r / g.my;
    T s = 0;
#pragma unroll
    for (int q = 0; q < NQ; q++) s += coef[q * g.n + t] * (T)x[nbr_idx(g, k, j, i, q)];
    y[t] = b ? b[t] - s : s;
}

// The same product without stored coefficients, for grids with no blanked
// point: each row's 19 coefficients are rebuilt from the nine precomputed
// face metric products of its own point and of its i-1, j-1 and k-1
// neighbours, in the assembly's operation order, and summed in the stored
// path's slot order, so the result has the same bits as k_apply on the
// assembled coefficients (the row mask drops the ghost columns as there).
template <typename T, typename X = T>
__global__ void __launch_bounds__(PTPB) k_apply_mf(Grid g, FaceG F, double th, const unsigned char *__restrict__ row,
                                                   const X *__restrict__ x, const T *__restrict__ b, T *__restrict__ y)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= g.n) return;
    if (!row[t]) { y[t] = 0; return; }
    const int t32 = (int)t; int i = t32 % g.mx; int r = t32 / g.mx; int j = r % g.my, k = r / g.my;
    double vol[NST];
    assemble_row<true>(g, F, nullptr, th, k, j, i, vol);
    T s = 0;
    if (i >= 2 && i <= g.mx - 3 && j >= 2 && j <= g.my - 3 && k >= 2 && k <= g.mz - 3) {
        // every neighbour is an interior point, hence a row (no point is
        // blanked on this path): no wrap and no mask reads
#pragma unroll
        for (int q = 0; q < NST; q++)
            s += (T)vol[q] * (T)x[t + ((long long)soff(q, 2) * g.my + soff(q, 1)) * g.mx + soff(q, 0)];
    } else {
#pragma unroll
        for (int q = 0; q < NST; q++) {
            const long long nb = nbr_idx(g, k, j, i, q);
            const T c = (q == 0 || row[nb]) ? (T)vol[q] : (T)0.;
            s += c * (T)x[nb];
        }
    }
    y[t] = b ? b[t] - s : s;
}

// ---- zebra line Gauss-Seidel over wall-normal (j) lines --------------------
// Lines (i,k) are coloured by (i+k) & 1.  A sweep relaxes colour 0 then colour
// 1; a half-sweep folds the off-line part of every row into a line
// right-hand side (reading the other colour's current values), then runs the
// forward and backward Thomas recurrences of each line with factors
// precomputed at setup.  The full and FP64 forms do this in two kernels (a
// fully parallel right-hand-side pass, then the line solve); the compact form
// in one (k_zebra_line_c4).  The smoother's coefficients live in a
// colour-split copy so that the passes read contiguous memory: plane
// [slot][colour][k][j][h] with i = i0 + 2h.
constexpr int NOFF = 16;
__host__ __device__ constexpr int offline_slot(int u)
{
    return u == 0 ? EP : u == 1 ? WP : u == 2 ? TP : u == 3 ? BP : u == 4 ? NE : u == 5 ? SE : u == 6 ? NW : u == 7 ? SW :
           u == 8 ? TN : u == 9 ? BN : u == 10 ? TS : u == 11 ? BS : u == 12 ? TE : u == 13 ? BE : u == 14 ? TW : BW;
}
enum { SM_LO = NOFF, SM_M, SM_CP, NSM };

struct ZGrid {
    int nkI, njI, nh;      // interior k and j counts, lines per colour per k row
    long long cplane;      // points per colour = nkI * njI * nh
};

__host__ __device__ inline int zebra_i0(int k, int c) { return 1 + ((1 + k + c) & 1); }

// Colour-split ("S") layout of a level vector: the interior points only, in
// the smoother's order [colour][k-1][j-1][h] with i = zebra_i0(k, colour) + 2h
// (the colour of (i, k) is (i + k) & 1).  The finest level's iterate and
// right-hand side take it on the compact one-pass path, so that a half-sweep
// reads and writes whole cache lines of one colour.
__host__ __device__ inline bool zs_interior(const Grid &g, int k, int j, int i)
{
    return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
__host__ __device__ inline long long zs_index(const ZGrid &z, int k, int j, int i)
{
    const int c = (i + k) & 1;
    const int h = (i - zebra_i0(k, c)) >> 1;
    return c * z.cplane + ((long long)(k - 1) * z.njI + (j - 1)) * z.nh + h;
}
// the same in 32-bit arithmetic (grids hold fewer than 2^31 points, so every
// S position does too), for the per-point kernels
__device__ inline unsigned zs_index32(const ZGrid &z, int k, int j, int i)
{
    const unsigned c = (unsigned)(i + k) & 1u;
    const unsigned h = (unsigned)(i - zebra_i0(k, (int)c)) >> 1;
    return c * (unsigned)z.cplane + ((unsigned)(k - 1) * (unsigned)z.njI + (unsigned)(j - 1)) * (unsigned)z.nh + h;
}
// (k, j, i) of natural point t of g, in 32-bit arithmetic
__device__ inline void nat_decode(const Grid &g, unsigned t, int &k, int &j, int &i)
{
    const unsigned mx = (unsigned)g.mx, my = (unsigned)g.my;
    const unsigned r = t / mx;
    i = (int)(t - r * mx);
    k = (int)(r / my);
    j = (int)(r - (unsigned)k * my);
}

// The FP32 smoother keeps the 16 off-line coefficients of a row divided by
// its diagonal, as FP16 pairs (8 planes of 32-bit words, slots 2p and 2p+1
// of offline_slot), with the diagonal in an FP32 plane: 36 bytes per point
// instead of 64, for the pass that reads them on every sweep.  The ratios are
// O(1) for the Poisson stencil and FP16 keeps 11 bits of them; the smoother
// is a preconditioner of the FP64 FGMRES, whose residual stays exact.
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
// Compact FP32 V-cycle level, for levels whose cross couplings (the twelve
// off-line neighbours off the i and k axes: edges through j and the i-k
// diagonals) are weak: the level works on the axis part of its operator.
// Its smoother keeps the ratios of the four axis neighbours (i+1, i-1, k+1,
// k-1 in the line's own row) as one 8-byte word, and its residual reads the
// centre and the six axis coefficients (7 of the 19 FP32 planes).
typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
// Cross-coupling strength (max over rows of sum |a_cross| / |a_P|) up to
// which a level takes the compact form.  On a 128x96x128 channel whose nodes
// were displaced by short-wavelength sines of growing amplitude the compact
// smoother took as many FGMRES iterations as the full one at strength 0.0085
// (7 % less solve time; 11 % with the axis-only residual too), one more per
// solve from 0.033 (slower), 8-9 instead of 6 at 0.083.  A channel whose only
// distortion spans the whole box (2 % of the spacing) measures 0.0012.
constexpr double COMPACT_MAX_CROSS = 0.01;

// a / d as a float within the FP16 range (the clamp only guards degenerate rows)
__device__ inline float half_ratio(double a, double d)
{
    const float v = (float)(a / d);
    return v > 65504.f ? 65504.f : (v < -65504.f ? -65504.f : v);
}

// Line right-hand sides of the FP32 smoother, rz = b - d * sum (a/d) x over
// the off-line points with the FP16 ratios, for two consecutive points
// (j, j+1) of a line per thread: their off-line neighbours share the rows j
// and j+1, so a pair reads 24 values of x instead of 32.
__global__ void __launch_bounds__(PTPB) k_zebra_rhs_h2(Grid g, ZGrid z, int c, const h2_t *__restrict__ sh,
                                                       const float *__restrict__ sd, const float *__restrict__ x,
                                                       const float *__restrict__ b, float *__restrict__ rz)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int npair = (z.njI + 1) >> 1;
    if (t >= (long long)z.nkI * npair * z.nh) return;
    const int h = (int)t % z.nh;
    const int r = (int)t / z.nh;
    const int jp = r % npair, kk = r / npair;
    const int k = kk + 1, j = 2 * jp + 1, i = zebra_i0(k, c) + 2 * h;
    if (i > g.mx - 2) return;
    const int np = j + 1 <= g.my - 2 ? 2 : 1;
    const int I[3] = {canon(i - 1, g.mx, g.pi), i, canon(i + 1, g.mx, g.pi)};
    const int K[3] = {canon(k - 1, g.mz, g.pk), k, canon(k + 1, g.mz, g.pk)};
    const int J[4] = {canon(j - 1, g.my, g.pj), j, canon(j + 1, g.my, g.pj), np == 2 ? canon(j + 2, g.my, g.pj) : j};
    const long long sstride = 2 * z.cplane;
    const long long cl0 = ((long long)kk * z.njI + 2 * jp) * z.nh + h;  // colour-local index of point j
#pragma unroll
    for (int p = 0; p < 2; p++) {
        if (p >= np) break;
        const long long cl = cl0 + p * z.nh, cq = c * z.cplane + cl;
        h2_t w[NOFF / 2];
#pragma unroll
        for (int u = 0; u < NOFF / 2; u++) w[u] = sh[u * sstride + cq];
        float acc = 0;
#pragma unroll
        for (int u = 0; u < NOFF; u++) {
            const int sl = offline_slot(u);
            const float xv = x[((long long)K[soff(sl, 2) + 1] * g.my + J[p + soff(sl, 1) + 1]) * g.mx + I[soff(sl, 0) + 1]];
            acc += (float)((u & 1) ? w[u / 2].y : w[u / 2].x) * xv;
        }
        rz[cl] = b[((long long)k * g.my + j + p) * g.mx + i] - sd[cq] * acc;
    }
}

// Compact line right-hand side of one point: b - d (r_E x_E + r_W x_W +
// r_T x_T + r_B x_B), the neighbours in the point's own row j
__device__ inline float c4_sum(h4_t w, float d, float xe, float xw, float xt, float xb, float bq)
{
    float acc = 0;
    acc += (float)w.x * xe;
    acc += (float)w.y * xw;
    acc += (float)w.z * xt;
    acc += (float)w.w * xb;
    return bq - d * acc;
}
__device__ inline float c4_rhs(const Grid &g, const int I[3], const int K[3], int j, h4_t w, float d, const float *__restrict__ x,
                               float bq)
{
    const long long row0 = (long long)j * g.mx;
    const float xe = x[(long long)K[1] * g.plane + row0 + I[2]];
    const float xw = x[(long long)K[1] * g.plane + row0 + I[0]];
    const float xt = x[(long long)K[2] * g.plane + row0 + I[1]];
    const float xb = x[(long long)K[0] * g.plane + row0 + I[1]];
    return c4_sum(w, d, xe, xw, xt, xb, bq);
}

// compact line right-hand sides of colour c, one thread per line point
__global__ void __launch_bounds__(PTPB) k_zebra_rhs_c4(Grid g, ZGrid z, int c, const h4_t *__restrict__ sh,
                                                       const float *__restrict__ sd, const float *__restrict__ x,
                                                       const float *__restrict__ b, float *__restrict__ rz)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= z.cplane) return;
    const int h = (int)t % z.nh;
    const int r = (int)t / z.nh;
    const int jj = r % z.njI, kk = r / z.njI;
    const int k = kk + 1, j = jj + 1, i = zebra_i0(k, c) + 2 * h;
    if (i > g.mx - 2) return;
    const int I[3] = {canon(i - 1, g.mx, g.pi), i, canon(i + 1, g.mx, g.pi)};
    const int K[3] = {canon(k - 1, g.mz, g.pk), k, canon(k + 1, g.mz, g.pk)};
    const long long cq = c * z.cplane + t;
    rz[t] = c4_rhs(g, I, K, j, sh[cq], sd[cq], x, b[((long long)k * g.my + j) * g.mx + i]);
}

// colour-split off-line coefficients and Thomas factors (lo, 1/pivot,
// upper/pivot).  T = double: sc holds the 16 off-line planes, then the three
// factor planes.  T = float: sc holds the three factor planes, sh the FP16
// coefficient pairs (compact: one h4_t word of the four axis ratios per
// point), sd the diagonal.
template <typename T>
__global__ void k_zebra_setup(Grid g, ZGrid z, const double *__restrict__ c64, const unsigned char *__restrict__ row,
                              T *__restrict__ sc, h2_t *__restrict__ sh, float *__restrict__ sd, int compact)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= 2LL * z.nkI * z.nh) return;
    const int h = (int)t % z.nh;
    const int r = (int)t / z.nh;
    const int kk = r % z.nkI, c = r / z.nkI, k = kk + 1;
    const int i = zebra_i0(k, c) + 2 * h;
    const bool valid = i <= g.mx - 2;
    const long long sstride = 2 * z.cplane;
    double cpp = 0;
    for (int jj = 0; jj < z.njI; jj++) {
        const long long cq = c * z.cplane + ((long long)kk * z.njI + jj) * z.nh + h;
        const long long q = valid ? ((long long)k * g.my + jj + 1) * g.mx + i : 0;
        const bool rw = valid && row[q];
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int u = 0; u < NOFF; u++) sc[u * sstride + cq] = rw ? (T)c64[offline_slot(u) * g.n + q] : T(0);
        } else {
            const double dg = rw ? c64[CP * g.n + q] : 1.0;
            if (compact) {
                h4_t w = {(_Float16)0.f, (_Float16)0.f, (_Float16)0.f, (_Float16)0.f};
                if (rw) {
                    w.x = (_Float16)half_ratio(c64[offline_slot(0) * g.n + q], dg);
                    w.y = (_Float16)half_ratio(c64[offline_slot(1) * g.n + q], dg);
                    w.z = (_Float16)half_ratio(c64[offline_slot(2) * g.n + q], dg);
                    w.w = (_Float16)half_ratio(c64[offline_slot(3) * g.n + q], dg);
                }
                reinterpret_cast<h4_t *>(sh)[cq] = w;
            } else {
#pragma unroll
                for (int u = 0; u < NOFF; u += 2) {
                    h2_t w = {(_Float16)0.f, (_Float16)0.f};
                    if (rw) {
                        w.x = (_Float16)half_ratio(c64[offline_slot(u) * g.n + q], dg);
                        w.y = (_Float16)half_ratio(c64[offline_slot(u + 1) * g.n + q], dg);
                    }
                    sh[(u / 2) * sstride + cq] = w;
                }
            }
            sd[cq] = rw ? (float)dg : 0.f;
        }
        T *sf = sizeof(T) == 8 ? sc + SM_LO * sstride : sc;  // lo, m, cp planes
        double lo = 0, m = 0, cpn = 0;
        if (rw) {
            lo = c64[SP * g.n + q];
            m = 1.0 / (c64[CP * g.n + q] - lo * cpp);
            cpn = c64[NP * g.n + q] * m;
        }
        sf[0 * sstride + cq] = (T)lo;
        sf[1 * sstride + cq] = (T)m;
        sf[2 * sstride + cq] = (T)cpn;
        cpp = cpn;
    }
}

// line right-hand sides of colour c: rz = b - (off-line part of A) x
template <typename T>
__global__ void __launch_bounds__(PTPB) k_zebra_rhs(Grid g, ZGrid z, int c, const T *__restrict__ sc,
                                                    const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ rz)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= z.cplane) return;
    const int h = (int)t % z.nh;
    const int r = (int)t / z.nh;
    const int jj = r % z.njI, kk = r / z.njI;
    const int k = kk + 1, j = jj + 1, i = zebra_i0(k, c) + 2 * h;
    if (i > g.mx - 2) return;
    const long long q = ((long long)k * g.my + j) * g.mx + i;
    const T *s0 = sc + c * z.cplane + t;
    const long long sstride = 2 * z.cplane;
    T acc = b[q];
#pragma unroll
    for (int u = 0; u < NOFF; u++) acc -= s0[u * sstride] * x[nbr_idx(g, k, j, i, offline_slot(u))];
    rz[t] = acc;
}


// Thomas recurrences of colour c, one thread per line; rz holds the line
// right-hand sides on entry and the forward-eliminated values afterwards.
// Loads are issued U points ahead of the dependent chain.
template <typename T, int U>
__global__ void __launch_bounds__(PTPB) k_zebra_solve(Grid g, ZGrid z, int c, const T *__restrict__ sf,
                                                      T *__restrict__ rz, T *__restrict__ x)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)z.nkI * z.nh) return;
    const int h = t % z.nh, kk = t / z.nh, k = kk + 1;
    const int i = zebra_i0(k, c) + 2 * h;
    if (i > g.mx - 2) return;
    const long long sstride = 2 * z.cplane;
    const long long base = (long long)kk * z.njI * z.nh + h;
    const T *lo = sf + c * z.cplane + base;
    const T *mm = sf + sstride + c * z.cplane + base;
    const T *cp = sf + 2 * sstride + c * z.cplane + base;
    T *rr = rz + base;
    const int nh = z.nh, nj = z.njI;
    T d = 0;
    for (int j0 = 0; j0 < nj; j0 += U) {
        T a[U], m[U], rv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int jj = j0 + u;
            if (jj < nj) { a[u] = lo[(long long)jj * nh]; m[u] = mm[(long long)jj * nh]; rv[u] = rr[(long long)jj * nh]; }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int jj = j0 + u;
            if (jj < nj) {
                d = (rv[u] - a[u] * d) * m[u];
                rr[(long long)jj * nh] = d;
            }
        }
    }
    const long long xrow = (long long)g.mx, xk = ((long long)k * g.my + 1) * g.mx + i;
    T xl = 0;
    for (int j0 = nj - 1; j0 >= 0; j0 -= U) {
        T cc[U], dd[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int jj = j0 - u;
            if (jj >= 0) { cc[u] = cp[(long long)jj * nh]; dd[u] = rr[(long long)jj * nh]; }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int jj = j0 - u;
            if (jj >= 0) {
                xl = dd[u] - cc[u] * xl;
                x[xk + jj * xrow] = xl;
            }
        }
    }
}

// Line solve of colour c with the line right-hand sides rz (from
// k_zebra_rhs).  A block owns 32 lines; each line is cut into nseg segments of
// at most SLM points, one thread per (line, segment), so the whole solve runs
// with one thread per ~SLM line points instead of one per line.  The Thomas
// forward recurrence d_j = (r_j - lo_j d_{j-1}) m_j is affine in d_{j-1}: each
// thread composes its segment's map (d_out = A d_in + B), the maps are chained
// across segments through LDS, and each thread then reruns the recurrence
// from its own d_in; the backward recurrence x_j = d_j - cp_j x_{j+1} is done
// the same way from the top segment down.  The segment's values stay in
// registers between the passes.
constexpr int LINE_LINES = 32, LINE_MAXSEG = 32;
// points per segment: the segment's right-hand sides and its three Thomas
// factors are held in registers (4 x SLM values).  FP32 lines of up to
// 32 x 8 points take segments of 8 (0.15 ms less per step than 16 at
// 256x192x256: more, lighter threads), longer ones 16.
template <typename T>
constexpr int line_slm()
{
    return sizeof(T) == 4 ? 16 : 8;
}
constexpr int LINE_SLM_SHORT = 8;

template <typename T, int SLM>
__global__ void __launch_bounds__(LINE_LINES * LINE_MAXSEG) k_zebra_scan(Grid g, ZGrid z, int c, int nseg, int sl,
                                                                        const T *__restrict__ sf, const T *__restrict__ rz,
                                                                        T *__restrict__ x)
{
    __shared__ T sA[LINE_MAXSEG * LINE_LINES], sB[LINE_MAXSEG * LINE_LINES];
    const int ln = threadIdx.x % LINE_LINES, seg = threadIdx.x / LINE_LINES;
    const long long nlines = (long long)z.nkI * z.nh;
    const long long lid = (long long)blockIdx.x * LINE_LINES + ln;
    int kk = 0, h = 0, k = 1, i = 1;
    bool act = lid < nlines;
    if (act) {
        kk = (int)(lid / z.nh);
        h = (int)(lid % z.nh);
        k = kk + 1;
        i = zebra_i0(k, c) + 2 * h;
        act = i <= g.mx - 2;
    }
    const int j0 = seg * sl;
    const int j1 = j0 + sl < z.njI ? j0 + sl : z.njI;
    const long long sstride = 2 * z.cplane;
    const long long base = (long long)kk * z.njI * z.nh + h;  // + jj * nh
    const T *s0 = sf + c * z.cplane + base;
    const T *lo = s0, *mm = s0 + sstride, *cp = s0 + 2 * sstride;
    const T *rr = rz + base;
    const long long nh = z.nh;
    T r[SLM], l[SLM], m[SLM], u[SLM];
#pragma unroll
    for (int v = 0; v < SLM; v++) {
        const int jj = j0 + v;
        if (act && jj < j1) {
            r[v] = rr[jj * nh];
            l[v] = lo[jj * nh];
            m[v] = mm[jj * nh];
            u[v] = cp[jj * nh];
        } else {
            r[v] = 0; l[v] = 0; m[v] = 0; u[v] = 0;
        }
    }
    // forward map of the segment: d_out = A d_in + B
    T A = 1, B = 0;
#pragma unroll
    for (int v = 0; v < SLM; v++)
        if (j0 + v < j1) {
            B = (r[v] - l[v] * B) * m[v];
            A = (-l[v] * A) * m[v];
        }
    const int me = seg * LINE_LINES + ln;
    sA[me] = A;
    sB[me] = B;
    __syncthreads();
    T d = 0;
    for (int s = 0; s < seg; s++) d = sA[s * LINE_LINES + ln] * d + sB[s * LINE_LINES + ln];
#pragma unroll
    for (int v = 0; v < SLM; v++)
        if (j0 + v < j1) {
            d = (r[v] - l[v] * d) * m[v];
            r[v] = d;
        }
    __syncthreads();
    // backward map of the segment: x_first = A x_above + B
    A = 1;
    B = 0;
#pragma unroll
    for (int v = SLM - 1; v >= 0; v--)
        if (j0 + v < j1) {
            B = r[v] - u[v] * B;
            A = -u[v] * A;
        }
    sA[me] = A;
    sB[me] = B;
    __syncthreads();
    T xl = 0;
    for (int s = nseg - 1; s > seg; s--) xl = sA[s * LINE_LINES + ln] * xl + sB[s * LINE_LINES + ln];
    if (act) {
#pragma unroll
        for (int v = SLM - 1; v >= 0; v--) {
            const int jj = j0 + v;
            if (jj < j1) {
                xl = r[v] - u[v] * xl;
                x[((long long)k * g.my + jj + 1) * g.mx + i] = xl;
            }
        }
    }
}

// Compact line relaxation of colour c in one pass: each (line, segment) thread
// forms its points' line right-hand sides as k_zebra_rhs_c4 does, in
// registers, and solves them as k_zebra_scan does, so rz is neither written
// nor read back and a half-sweep is one launch instead of two.  The compact
// form's off-line neighbours (i+-1 and k+-1 in the line's own row) all belong
// to the other colour when every periodic direction has an even interior
// count (zebra_fused_ok), and this pass writes only colour c, so the values it
// reads are the ones the two-pass form reads: the same bits.  Blocks take
// their 32 lines in XCD order, so the k+-1 rows a block reads were read by the
// blocks just before it on the same XCD.
__host__ __device__ inline bool zebra_fused_ok(const Grid &g)
{
    return (!g.pi || ((g.mx - 2) & 1) == 0) && (!g.pk || ((g.mz - 2) & 1) == 0);
}

// LL lines per block (32, or 64 with at most 16 segments per line)
template <int SLM, bool XZ, bool SPL, int LL>
__global__ void __launch_bounds__(1024) k_zebra_line_c4(Grid g, ZGrid z, int c, int nseg, int sl,
                                                                           const h4_t *__restrict__ sh,
                                                                           const float *__restrict__ sd,
                                                                           const float *__restrict__ b,
                                                                           const float *__restrict__ sf, float *x)
{
    static_assert(LL * (LL == 64 ? 16 : LINE_MAXSEG) <= 1024, "threads per block");
    __shared__ float sA[1024], sB[1024];
    const int ln = threadIdx.x % LL, seg = threadIdx.x / LL;
    const long long nlines = (long long)z.nkI * z.nh;
    const long long lid = (long long)xcd_block1() * LL + ln;
    int kk = 0, h = 0, k = 1, i = 1;
    bool act = lid < nlines;
    if (act) {
        kk = (int)(lid / z.nh);
        h = (int)(lid % z.nh);
        k = kk + 1;
        i = zebra_i0(k, c) + 2 * h;
        act = i <= g.mx - 2;
    }
    const int j0 = seg * sl;
    const int j1 = j0 + sl < z.njI ? j0 + sl : z.njI;
    const long long sstride = 2 * z.cplane;
    const long long base = (long long)kk * z.njI * z.nh + h;  // + jj * nh
    const long long cb = c * z.cplane + base;
    const float *lo = sf + cb, *mm = sf + sstride + cb, *cp = sf + 2 * sstride + cb;
    const long long nh = z.nh;
    const int I[3] = {canon(i - 1, g.mx, g.pi), i, canon(i + 1, g.mx, g.pi)};
    const int K[3] = {canon(k - 1, g.mz, g.pk), k, canon(k + 1, g.mz, g.pk)};
    // SPL: b and x in the S layout.  The neighbours are of the other colour
    // (q): rows kk (i +- 1), K[0] - 1 and K[2] - 1 (i, at the line's own h);
    // a non-periodic ghost neighbour reads as the zero it holds in the
    // natural layout
    const int q = 1 - c;
    const bool okw = I[0] >= 1 && I[0] <= g.mx - 2, oke = I[2] >= 1 && I[2] <= g.mx - 2;
    const bool okb = K[0] >= 1 && K[0] <= g.mz - 2, okt = K[2] >= 1 && K[2] <= g.mz - 2;
    const long long hw = okw ? (I[0] - zebra_i0(k, q)) >> 1 : 0, he = oke ? (I[2] - zebra_i0(k, q)) >> 1 : 0;
    const long long rowq = q * z.cplane, nrow = (long long)z.njI * z.nh;
    const long long sw = rowq + kk * nrow + hw, se = rowq + kk * nrow + he;
    const long long sbt = rowq + (okb ? K[0] - 1 : 0) * nrow + h, stp = rowq + (okt ? K[2] - 1 : 0) * nrow + h;
    float r[SLM], l[SLM], m[SLM], u[SLM];
#pragma unroll
    for (int v = 0; v < SLM; v++) {
        const int jj = j0 + v;
        if (act && jj < j1) {
            const long long o = jj * nh;
            const float bq = SPL ? b[cb + o] : b[((long long)k * g.my + jj + 1) * g.mx + i];
            // XZ: the iterate is zero on entry (the first half-sweep of a
            // V-cycle level): the same sum over zero neighbours, without reading them
            if constexpr (XZ)
                r[v] = c4_sum(sh[cb + o], sd[cb + o], 0.f, 0.f, 0.f, 0.f, bq);
            else if constexpr (SPL)
                r[v] = c4_sum(sh[cb + o], sd[cb + o], oke ? x[se + o] : 0.f, okw ? x[sw + o] : 0.f, okt ? x[stp + o] : 0.f,
                              okb ? x[sbt + o] : 0.f, bq);
            else
                r[v] = c4_rhs(g, I, K, jj + 1, sh[cb + o], sd[cb + o], x, bq);
            l[v] = lo[o];
            m[v] = mm[o];
            u[v] = cp[o];
        } else {
            r[v] = 0; l[v] = 0; m[v] = 0; u[v] = 0;
        }
    }
    // forward map of the segment: d_out = A d_in + B
    float A = 1, B = 0;
#pragma unroll
    for (int v = 0; v < SLM; v++)
        if (j0 + v < j1) {
            B = (r[v] - l[v] * B) * m[v];
            A = (-l[v] * A) * m[v];
        }
    const int me = seg * LL + ln;
    sA[me] = A;
    sB[me] = B;
    __syncthreads();
    float d = 0;
    for (int s = 0; s < seg; s++) d = sA[s * LL + ln] * d + sB[s * LL + ln];
#pragma unroll
    for (int v = 0; v < SLM; v++)
        if (j0 + v < j1) {
            d = (r[v] - l[v] * d) * m[v];
            r[v] = d;
        }
    __syncthreads();
    // backward map of the segment: x_first = A x_above + B
    A = 1;
    B = 0;
#pragma unroll
    for (int v = SLM - 1; v >= 0; v--)
        if (j0 + v < j1) {
            B = r[v] - u[v] * B;
            A = -u[v] * A;
        }
    sA[me] = A;
    sB[me] = B;
    __syncthreads();
    float xl = 0;
    for (int s = nseg - 1; s > seg; s--) xl = sA[s * LL + ln] * xl + sB[s * LL + ln];
    if (act) {
#pragma unroll
        for (int v = SLM - 1; v >= 0; v--) {
            const int jj = j0 + v;
            if (jj < j1) {
                xl = r[v] - u[v] * xl;
                if (SPL) x[cb + jj * nh] = xl;
                else x[((long long)k * g.my + jj + 1) * g.mx + i] = xl;
            }
        }
    }
}

// coarsest level: x = B^{-1} b over the rows, B^{-1} stored column-major.
// Each block gathers b into LDS once.
constexpr int DENSE_ROWS = 64, DENSE_SPLIT = 16, DENSE_LDS = 1024;
// x = Inv b on the coarsest level: a block takes DENSE_ROWS rows; its
// DENSE_SPLIT groups of DENSE_ROWS threads each sum one contiguous slice of the
// columns (the inverse is column-major, so a group's loads are coalesced), and
// the slices are added in group order (a fixed order).  One thread per row
// summing all columns serially was latency-bound (29.5 us for 384 rows).
template <typename T>
__global__ void __launch_bounds__(DENSE_ROWS * DENSE_SPLIT) k_dense_solve(int nr, const int *__restrict__ ridx,
                                                                          const T *__restrict__ inv, const T *__restrict__ b,
                                                                          T *__restrict__ x)
{
    __shared__ T bs[DENSE_LDS];
    __shared__ T part[DENSE_SPLIT][DENSE_ROWS];
    for (int c = threadIdx.x; c < nr; c += blockDim.x) bs[c] = b[ridx[c]];
    __syncthreads();
    const int lr = threadIdx.x % DENSE_ROWS, grp = threadIdx.x / DENSE_ROWS;
    const int r = blockIdx.x * DENSE_ROWS + lr;
    const int chunk = (nr + DENSE_SPLIT - 1) / DENSE_SPLIT;
    const int c0 = grp * chunk, c1 = c0 + chunk < nr ? c0 + chunk : nr;
    T s = 0;
    if (r < nr)
        for (int c = c0; c < c1; c++) s += inv[(long long)c * nr + r] * bs[c];
    part[grp][lr] = s;
    __syncthreads();
    if (grp == 0 && r < nr) {
        T t = part[0][lr];
        for (int g = 1; g < DENSE_SPLIT; g++) t += part[g][lr];
        x[ridx[r]] = t;
    }
}

// coarse b = sum of the eight child residuals
template <typename T>
__global__ void k_restrict(Grid gc, Grid gf, const unsigned char *__restrict__ rowc, const unsigned char *__restrict__ rowf,
                           const T *__restrict__ rf, T *__restrict__ bc)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= gc.n) return;
    if (!rowc[t]) { bc[t] = 0; return; }
    const int t32 = (int)t; int i = t32 % gc.mx; int r = t32 / gc.mx; int j = r % gc.my, k = r / gc.my;
    T s = 0;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                long long qf = gidx(gf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di);
                if (rowf[qf]) s += rf[qf];
            }
    bc[t] = s;
}

// k_apply's residual r = b - A x (its first NQ stencil slots) at the eight
// children of each coarse point, summed into the coarse b as k_restrict sums
// them: the fine residual is never stored.  Each child's residual is formed
// with k_apply's expression in its slot order and the children are added in
// k_restrict's order, so bc has the same bits as the two-kernel form.  The
// x values of the 4x4x4 fine block around the children come through the
// per-axis canonical positions nbr_idx uses.
// SPL: the fine x and b in the S layout (zf), a ghost x reading as zero.
template <typename T, int NQ, bool SPL = false>
__global__ void __launch_bounds__(PTPB) k_residual_restrict(Grid gc, Grid gf, const T *__restrict__ coef,
                                                            const unsigned char *__restrict__ rowc,
                                                            const unsigned char *__restrict__ rowf, const T *__restrict__ x,
                                                            const T *__restrict__ b, T *__restrict__ bc, ZGrid zf = {})
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= gc.n) return;
    if (!rowc[t]) { bc[t] = 0; return; }
    const int t32 = (int)t; int i = t32 % gc.mx; int r = t32 / gc.mx; int j = r % gc.my, k = r / gc.my;
    // fine positions 2c-2 .. 2c+1 per axis (children at 2c-1 and 2c)
    int I[4], J[4], K[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        I[d] = canon(2 * i - 2 + d, gf.mx, gf.pi);
        J[d] = canon(2 * j - 2 + d, gf.my, gf.pj);
        K[d] = canon(2 * k - 2 + d, gf.mz, gf.pk);
    }
    T s = 0;
#pragma unroll
    for (int dk = 0; dk < 2; dk++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
            for (int di = 0; di < 2; di++) {
                const long long qf = gidx(gf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di);
                if (!rowf[qf]) continue;
                T a = 0;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int kq = K[dk + 1 + soff(q, 2)], jq = J[dj + 1 + soff(q, 1)], iq = I[di + 1 + soff(q, 0)];
                    T xv;
                    if constexpr (SPL) xv = zs_interior(gf, kq, jq, iq) ? x[zs_index32(zf, kq, jq, iq)] : T(0);
                    else xv = x[((long long)kq * gf.my + jq) * gf.mx + iq];
                    a += coef[q * gf.n + qf] * xv;
                }
                const T bq = SPL ? b[zs_index32(zf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di)] : b[qf];
                s += bq - a;
            }
    bc[t] = s;
}

__device__ inline int parent_nb(int I, int d, int mc, int per)
{
    int J = I + d;
    if (J < 1) J = per ? mc - 2 : 1;
    else if (J > mc - 2) J = per ? 1 : mc - 2;
    return J;
}

// fine x += trilinear(coarse x)
template <typename T>
__global__ void k_prolong(Grid gc, Grid gf, const unsigned char *__restrict__ rowf, const T *__restrict__ xc,
                          T *__restrict__ xf)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= gf.n) return;
    if (!rowf[t]) return;
    const int t32 = (int)t; int i = t32 % gf.mx; int r = t32 / gf.mx; int j = r % gf.my, k = r / gf.my;
    int I = (i + 1) / 2, J = (j + 1) / 2, K = (k + 1) / 2;
    int I2 = parent_nb(I, (i & 1) ? -1 : 1, gc.mx, gc.pi);
    int J2 = parent_nb(J, (j & 1) ? -1 : 1, gc.my, gc.pj);
    int K2 = parent_nb(K, (k & 1) ? -1 : 1, gc.mz, gc.pk);
#define XC(kk, jj, ii) xc[((long long)(kk) * gc.my + (jj)) * gc.mx + (ii)]
    T v = T(27) * XC(K, J, I) + T(9) * (XC(K, J, I2) + XC(K, J2, I) + XC(K2, J, I)) +
          T(3) * (XC(K, J2, I2) + XC(K2, J, I2) + XC(K2, J2, I)) + XC(K2, J2, I2);
#undef XC
    xf[t] += v / T(64);
}

// k_prolong on a fine iterate in the S layout: one thread per S position,
// the same interpolation expression
template <typename T>
__global__ void k_prolong_split(Grid gc, Grid gf, ZGrid zf, const unsigned char *__restrict__ rowf, const T *__restrict__ xc,
                                T *__restrict__ xs)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned cp = (unsigned)zf.cplane, nh = (unsigned)zf.nh, nj = (unsigned)zf.njI;
    if (t >= 2 * cp) return;
    const unsigned col = t >= cp ? 1u : 0u;
    const unsigned rem = t - col * cp;
    const unsigned r = rem / nh, h = rem - r * nh;
    const unsigned kk = r / nj, jj = r - kk * nj;
    const int j = (int)jj + 1, k = (int)kk + 1;
    const int i = zebra_i0(k, (int)col) + 2 * (int)h;
    if (i > gf.mx - 2) return;
    if (!rowf[((long long)k * gf.my + j) * gf.mx + i]) return;
    int I = (i + 1) / 2, J = (j + 1) / 2, K = (k + 1) / 2;
    int I2 = parent_nb(I, (i & 1) ? -1 : 1, gc.mx, gc.pi);
    int J2 = parent_nb(J, (j & 1) ? -1 : 1, gc.my, gc.pj);
    int K2 = parent_nb(K, (k & 1) ? -1 : 1, gc.mz, gc.pk);
#define XC(kk, jj, ii) xc[((long long)(kk) * gc.my + (jj)) * gc.mx + (ii)]
    T v = T(27) * XC(K, J, I) + T(9) * (XC(K, J, I2) + XC(K, J2, I) + XC(K2, J, I)) +
          T(3) * (XC(K, J2, I2) + XC(K2, J, I2) + XC(K2, J2, I)) + XC(K2, J2, I2);
#undef XC
    xs[t] += v / T(64);
}

// natural FP64 v -> S-layout FP32 (the finest right-hand side)
__global__ void k_convert_split(Grid g, ZGrid z, const double *__restrict__ v, float *__restrict__ bs)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)g.n) return;
    int k, j, i;
    nat_decode(g, t, k, j, i);
    if (zs_interior(g, k, j, i)) bs[zs_index32(z, k, j, i)] = (float)v[t];
}

// S-layout FP32 iterate -> natural FP32 vector, zero at the ghosts (as the
// natural finest iterate holds them)
__global__ void k_split_to_natural(Grid g, ZGrid z, const float *__restrict__ xs, float *__restrict__ x)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)g.n) return;
    int k, j, i;
    nat_decode(g, t, k, j, i);
    x[t] = zs_interior(g, k, j, i) ? xs[zs_index32(z, k, j, i)] : 0.f;
}

// coarse node coordinates: every other fine node (init.c:402-455)
__global__ void k_coarse_coor(Grid gc, Grid gf, const double *__restrict__ cf, double *__restrict__ cc)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= gc.n) return;
    const int t32 = (int)t; int i = t32 % gc.mx; int r = t32 / gc.mx; int j = r % gc.my, k = r / gc.my;
    if (i >= gc.mx - 1 || j >= gc.my - 1 || k >= gc.mz - 1) {
        cc[3 * t] = cc[3 * t + 1] = cc[3 * t + 2] = 0;
        return;
    }
    long long s = ((long long)(2 * k) * gf.my + 2 * j) * gf.mx + 2 * i;
    cc[3 * t] = cf[3 * s]; cc[3 * t + 1] = cf[3 * s + 1]; cc[3 * t + 2] = cf[3 * s + 2];
}

// a coarse cell is blanked only when all eight children are
__global__ void k_coarse_nvert(Grid gc, Grid gf, const double *__restrict__ nf, double *__restrict__ nc)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= gc.n) return;
    const int t32 = (int)t; int i = t32 % gc.mx; int r = t32 / gc.mx; int j = r % gc.my, k = r / gc.my;
    if (i < 1 || i > gc.mx - 2 || j < 1 || j > gc.my - 2 || k < 1 || k > gc.mz - 2) { nc[t] = 0; return; }
    double mn = 1e30;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) mn = fmin(mn, nf[gidx(gf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di)]);
    nc[t] = mn;
}

// ---- vector kernels --------------------------------------------------------
template <typename A, typename B>
__global__ void k_convert(long long n, const A *__restrict__ x, B *__restrict__ y)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) y[t] = (B)x[t];
}

template <typename T>
__global__ void k_zero(long long n, T *__restrict__ x)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) x[t] = 0;
}

// y = a * x   (a read from device memory so no host round trip is needed)
__global__ void k_scale_dev(long long n, const double *__restrict__ a, int inv, const double *__restrict__ x,
                            double *__restrict__ y)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    double s = inv ? 1.0 / sqrt(a[0]) : a[0];
    if (t < n) y[t] = s * x[t];
}

// out = <w, w> of the second Gram-Schmidt pass when it ran (gate as in
// k_multiaxpy), else the first pass's *gww
__global__ void __launch_bounds__(RED_THREADS) k_finish_gated(const double *__restrict__ part, int n,
                                                              const double *__restrict__ gww, const double *__restrict__ gww0,
                                                              double *__restrict__ out)
{
    if (!(gww[0] < 0.25 * gww0[0])) {
        if (threadIdx.x == 0) out[0] = gww[0];
        return;
    }
    double v = OpSum::init();
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = OpSum::op(v, part[q]);
    v = block_reduce<OpSum>(v);
    if (threadIdx.x == 0) out[0] = v;
}

// y = x / sqrt(ww) (0 when ww <= 0), ww read from device memory
// y = x / sqrt(ww[0]) (0 when ww[0] <= 0); y32, when given, receives y as
// FP32 (the next FP32 V-cycle's right-hand side)
// (natural layout, or the S layout of zs when zs.cplane > 0: interior points only)
__global__ void k_scale_rnorm(long long n, const double *__restrict__ ww, const double *__restrict__ x,
                              double *__restrict__ y, float *__restrict__ y32, Grid g, ZGrid zs)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const double hn = sqrt(ww[0] > 0 ? ww[0] : 0);
    const double s = hn > 0 ? 1.0 / hn : 0.0;
    if (t < n) {
        const double v = s * x[t];
        y[t] = v;
        if (y32) {
            if (zs.cplane > 0) {
                int k, j, i;
                nat_decode(g, (unsigned)t, k, j, i);
                if (zs_interior(g, k, j, i)) y32[zs_index32(zs, k, j, i)] = (float)v;
            } else {
                y32[t] = (float)v;
            }
        }
    }
}

constexpr int MAXV = 33;

// x = x + sum_v y[v] Z_v (Z in FP64, or FP32 from the FP32 V-cycle)
template <typename Z_T>
__global__ void k_multiupdate(long long n, int nv, const double *__restrict__ y, const Z_T *__restrict__ Z, long long ld,
                              double *__restrict__ x)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = x[t];
    for (int v = 0; v < nv; v++) s += y[v] * (double)Z[v * ld + t];
    x[t] = s;
}

// sum and count of the rows of x (plain mean, poisson_hypre.c:202-263)
__global__ void k_row_sum(long long n, const unsigned char *__restrict__ row, const double *__restrict__ x,
                          double *__restrict__ part)
{
    double s = 0, c = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        if (row[t]) { s += x[t]; c += 1; }
    s = block_reduce<OpSum>(s);
    c = block_reduce<OpSum>(c);
    if (threadIdx.x == 0) { part[blockIdx.x] = s; part[RED_BLOCKS + blockIdx.x] = c; }
}

__global__ void k_sub_mean(long long n, const unsigned char *__restrict__ row, const double *__restrict__ sc,
                           double *__restrict__ x)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n || !row[t]) return;
    double val = -sc[0] / sc[1];
    x[t] += val;
}

}  // namespace vfs

// ---------------------------------------------------------------------------
// solver state
// ---------------------------------------------------------------------------

struct PLevel {
    vfs::Grid g;
    vfs::ZGrid z;
    double *coor = nullptr;
    double *met[17] = {nullptr};
    double *nvert = nullptr;
    unsigned char *row = nullptr;
    double *c64 = nullptr;
    float *c32 = nullptr;
    // nine face metric products per point (finest level, no point blanked):
    // the FP64 operator is then applied matrix-free (k_apply_mf)
    double *fg = nullptr;
    // colour-split smoother coefficients + Thomas factors, one-colour line rhs
    // (FP32: sc32 = factors, sh16 = FP16 coefficient ratios, sd32 = diagonal)
    float *sc32 = nullptr, *rz32 = nullptr, *sd32 = nullptr;
    vfs::h2_t *sh16 = nullptr;
    // FP32 smoother form: compact (four axis ratios in sh16's memory, as
    // h4_t) when the level's cross-coupling strength xs is small
    int compact = 0;
    double xs = 0;
    double *sc64 = nullptr, *rz64 = nullptr;
    // work vectors
    float *x32 = nullptr, *b32 = nullptr, *r32 = nullptr;
    // the finest level's FP32 iterate and right-hand side in the S layout
    // (the compact one-pass path; finest set on level 0 only)
    float *xs32 = nullptr, *bs32 = nullptr;
    bool finest = false;
    double *x64 = nullptr, *b64 = nullptr, *r64 = nullptr;
    // dense inverse on the coarsest level (nr rows), column-major
    int nr = 0;
    int *ridx = nullptr;
    float *inv32 = nullptr;
    double *inv64 = nullptr;
};

// a captured coarse sub-cycle (levels 1..nl-1 of one V-cycle) and the
// smoothing counts it was captured with
// Smoothing sweeps of a V-cycle: pre and post on the levels below
// deep_level, deep_pre and deep_post from deep_level on (deep_level 0: every
// level takes pre and post), coarse on a coarsest level without a dense solve.
struct Sweeps {
    int pre, post, coarse, deep_level, deep_pre, deep_post;
    int pre_at(int l) const { return deep_level > 0 && l >= deep_level ? deep_pre : pre; }
    int post_at(int l) const { return deep_level > 0 && l >= deep_level ? deep_post : post; }
    bool operator==(const Sweeps &o) const
    {
        return pre == o.pre && post == o.post && coarse == o.coarse && deep_level == o.deep_level && deep_pre == o.deep_pre &&
               deep_post == o.deep_post;
    }
};

struct SubGraph {
    hipGraphExec_t exec = nullptr;
    Sweeps key{-1, -1, -1, -1, -1, -1};
};

namespace vfs {
struct AmgDev;
int amg_build(const double *c64_host, const unsigned char *row_host, long long n_pts,
              long long (*nbr)(const void *g, long long p, int s), const void *grid, double theta, AmgDev **out);
int amg_apply(hipStream_t st, AmgDev *a, long long n_pts, const double *v, double *z);
void amg_free(AmgDev *a);
double amg_theta(const AmgDev *a);
}  // namespace vfs

struct vfs_poisson {
    vfs::AmgDev *amg = nullptr;  // algebraic multigrid hierarchy of level 0 (when requested)
    int nl = 0;          // levels built
    int req_levels = 0;  // levels requested when it was built
    int restart = 0;
    PLevel L[8];
    double *V = nullptr, *Z = nullptr, *w = nullptr, *dev_scalars = nullptr;
    std::vector<void *> allocs;
    SubGraph sub[2];  // [0] FP32, [1] FP64 hierarchy
};

namespace vfs {

// largest coarsest level solved by a dense inverse
constexpr int DENSE_MAX = 1024;
static_assert(DENSE_MAX <= DENSE_LDS, "k_dense_solve stages the whole coarse right-hand side in LDS");

// Every buffer starts zeroed: ghost cells and blanked cells of the work
// vectors are read (times a zero coefficient) but never written, and garbage
// there could hold NaN.
static int palloc(vfs_poisson *ps, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return check_hip(e, "hipMalloc (poisson)");
    ps->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) return check_hip(e, "hipMemset (poisson)");
    return VFS_OK;
}

void poisson_free(vfs_ctx *c)
{
    if (!c || !c->ps) return;
    hipStreamSynchronize(c->stream);
    for (SubGraph &s : c->ps->sub)
        if (s.exec) hipGraphExecDestroy(s.exec);
    for (void *p : c->ps->allocs) hipFree(p);
    amg_free(c->ps->amg);
    delete c->ps;
    c->ps = nullptr;
}

static int level_alloc_work(vfs_poisson *ps, PLevel &L)
{
    size_t n = (size_t)L.g.n;
    L.z.nkI = L.g.mz - 2;
    L.z.njI = L.g.my - 2;
    L.z.nh = (L.g.mx - 2 + 1) / 2;
    L.z.cplane = (long long)L.z.nkI * L.z.njI * L.z.nh;
    const size_t cp = (size_t)L.z.cplane;
    int rc = 0;
    rc = rc ? rc : palloc(ps, (void **)&L.x32, n * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.b32, n * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.r32, n * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.x64, n * 8);
    rc = rc ? rc : palloc(ps, (void **)&L.b64, n * 8);
    rc = rc ? rc : palloc(ps, (void **)&L.r64, n * 8);
    rc = rc ? rc : palloc(ps, (void **)&L.row, n);
    rc = rc ? rc : palloc(ps, (void **)&L.c32, NST * n * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.c64, NST * n * 8);
    rc = rc ? rc : palloc(ps, (void **)&L.sc32, 3 * 2 * cp * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.sh16, NOFF / 2 * 2 * cp * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.sd32, 2 * cp * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.sc64, NSM * 2 * cp * 8);
    rc = rc ? rc : palloc(ps, (void **)&L.rz32, cp * 4);
    rc = rc ? rc : palloc(ps, (void **)&L.rz64, cp * 8);
    return rc;
}

// cross-coupling strength of a level: max over its rows of sum |a_u| / |a_P|
// over the twelve off-line neighbours off the i and k axes
__global__ void __launch_bounds__(RED_THREADS) k_cross_strength(Grid g, const double *__restrict__ c64,
                                                                const unsigned char *__restrict__ row, double *__restrict__ part)
{
    double v = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < g.n; t += (long long)gridDim.x * blockDim.x) {
        if (!row[t]) continue;
        double s = 0;
        for (int u = 4; u < NOFF; u++) s += fabs(c64[offline_slot(u) * g.n + t]);
        const double d = fabs(c64[CP * g.n + t]);
        if (d > 0) v = fmax(v, s / d);
    }
    v = block_reduce<OpMax>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// The FP32 V-cycle's form for a level: compact when its cross couplings are
// weak (VFS_SMOOTH_COMPACT=0 or 1 forces the full or the compact form)
static int choose_smoother(vfs_ctx *c, PLevel &L)
{
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_cross_strength, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, L.g, L.c64, L.row, c->red);
    hipLaunchKernelGGL(finish_reduce<OpMax>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + RED_BLOCKS);
    VFS_LAUNCH_CHECK("cross-coupling strength");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    L.xs = c->red_host[0];
    const char *e = std::getenv("VFS_SMOOTH_COMPACT");
    if (e && e[0] == '0') L.compact = 0;
    else if (e && e[0] == '1') L.compact = 1;
    else L.compact = L.xs <= COMPACT_MAX_CROSS;
    const char *tr = std::getenv("VFS_SOLVER_TRACE");
    if (tr && tr[0] == '1')
        std::fprintf(stderr, "level %dx%dx%d: cross-coupling strength %.9g, %s FP32 smoother\n", L.g.mx, L.g.my, L.g.mz, L.xs,
                     L.compact ? "compact" : "full");
    return VFS_OK;
}

static int level_assemble(vfs_ctx *c, PLevel &L, double *const met[17], const double *nvert)
{
    hipStream_t st = c->stream;
    FaceSet F = {met[4], met[5], met[6], met[7], met[8], met[9], met[10], met[11], met[12], met[13], met[14], met[15]};
    const double th = c->p.poisson_threshold;
    hipLaunchKernelGGL(k_row_mask, dim3(blocks_for(L.g.n, PTPB)), dim3(PTPB), 0, st, L.g, nvert, th, L.row);
    hipLaunchKernelGGL(k_assemble, dim3(blocks_for(L.g.n, PTPB)), dim3(PTPB), 0, st, L.g, F, nvert, th, L.row, L.c64, L.c32);
    VFS_LAUNCH_CHECK("assemble");
    int rc = choose_smoother(c, L);
    if (rc) return rc;
    const long long lines = 2LL * L.z.nkI * L.z.nh;
    hipLaunchKernelGGL(k_zebra_setup<float>, dim3(blocks_for(lines, PTPB)), dim3(PTPB), 0, st, L.g, L.z, L.c64, L.row, L.sc32,
                       L.sh16, L.sd32, L.compact);
    hipLaunchKernelGGL(k_zebra_setup<double>, dim3(blocks_for(lines, PTPB)), dim3(PTPB), 0, st, L.g, L.z, L.c64, L.row, L.sc64,
                       nullptr, nullptr, 0);
    VFS_LAUNCH_CHECK("smoother setup");
    return VFS_OK;
}

// Dense inverse of B = A + s 1 1^T over the rows of the coarsest level.  A has
// the constant vector as its null space (rows are sums of face fluxes), so B
// is regular and B^{-1} b solves A x = b for every consistent b.
static int level_dense_inverse(vfs_ctx *c, vfs_poisson *ps, PLevel &L)
{
    const Grid &g = L.g;
    const size_t n = (size_t)g.n;
    std::vector<unsigned char> row(n);
    VFS_HIP(hipMemcpyAsync(row.data(), L.row, n, hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> map(n, -1), ridx;
    for (size_t q = 0; q < n; q++)
        if (row[q]) { map[q] = (int)ridx.size(); ridx.push_back((int)q); }
    const int nr = (int)ridx.size();
    if (nr == 0 || nr > DENSE_MAX) return VFS_OK;  // keep smoothing on this level
    std::vector<double> coef(NST * n);
    VFS_HIP(hipMemcpyAsync(coef.data(), L.c64, NST * n * 8, hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> A((size_t)nr * nr, 0.0), I((size_t)nr * nr, 0.0);
    double dmax = 0;
    for (int r = 0; r < nr; r++) {
        const long long q = ridx[r];
        const int i = q % g.mx, j = (q / g.mx) % g.my, k = q / ((long long)g.mx * g.my);
        for (int s = 0; s < NST; s++) {
            const double v = coef[(size_t)s * n + q];
            if (v == 0) continue;
            const int col = map[nbr_idx(g, k, j, i, s)];
            if (col >= 0) A[(size_t)r * nr + col] += v;
        }
        dmax = std::fmax(dmax, std::fabs(A[(size_t)r * nr + r]));
        I[(size_t)r * nr + r] = 1.0;
    }
    const double shift = dmax / nr;
    for (auto &v : A) v += shift;
    // Gauss-Jordan with partial pivoting: A -> identity, I -> B^{-1}
    for (int col = 0; col < nr; col++) {
        int piv = col;
        for (int r = col + 1; r < nr; r++)
            if (std::fabs(A[(size_t)r * nr + col]) > std::fabs(A[(size_t)piv * nr + col])) piv = r;
        if (A[(size_t)piv * nr + col] == 0) return set_error(VFS_EINVAL, "singular coarsest Poisson operator");
        if (piv != col)
            for (int q = 0; q < nr; q++) {
                std::swap(A[(size_t)piv * nr + q], A[(size_t)col * nr + q]);
                std::swap(I[(size_t)piv * nr + q], I[(size_t)col * nr + q]);
            }
        const double inv = 1.0 / A[(size_t)col * nr + col];
        for (int q = 0; q < nr; q++) { A[(size_t)col * nr + q] *= inv; I[(size_t)col * nr + q] *= inv; }
        for (int r = 0; r < nr; r++) {
            if (r == col) continue;
            const double f = A[(size_t)r * nr + col];
            if (f == 0) continue;
            for (int q = 0; q < nr; q++) {
                A[(size_t)r * nr + q] -= f * A[(size_t)col * nr + q];
                I[(size_t)r * nr + q] -= f * I[(size_t)col * nr + q];
            }
        }
    }
    // column-major copy: inv[c * nr + r] = B^{-1}[r][c]
    std::vector<double> t64((size_t)nr * nr);
    std::vector<float> t32((size_t)nr * nr);
    for (int r = 0; r < nr; r++)
        for (int q = 0; q < nr; q++) {
            t64[(size_t)q * nr + r] = I[(size_t)r * nr + q];
            t32[(size_t)q * nr + r] = (float)I[(size_t)r * nr + q];
        }
    int rc = palloc(ps, (void **)&L.ridx, (size_t)nr * sizeof(int));
    if (!rc) rc = palloc(ps, (void **)&L.inv64, (size_t)nr * nr * 8);
    if (!rc) rc = palloc(ps, (void **)&L.inv32, (size_t)nr * nr * 4);
    if (rc) return rc;
    VFS_HIP(hipMemcpyAsync(L.ridx, ridx.data(), (size_t)nr * sizeof(int), hipMemcpyHostToDevice, c->stream));
    VFS_HIP(hipMemcpyAsync(L.inv64, t64.data(), (size_t)nr * nr * 8, hipMemcpyHostToDevice, c->stream));
    VFS_HIP(hipMemcpyAsync(L.inv32, t32.data(), (size_t)nr * nr * 4, hipMemcpyHostToDevice, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    L.nr = nr;
    return VFS_OK;
}

static int poisson_setup(vfs_ctx *c, int levels)
{
    // the hierarchy is kept while the requested level count is the same (the
    // built count can be smaller: coarsening stops at an odd cell count)
    if (c->ps && c->ps->req_levels == levels) return VFS_OK;
    // the solver's point kernels decode (i, j, k) in 32-bit arithmetic
    if (c->g.n >= (1LL << 31)) return set_error(VFS_EINVAL, "Poisson solver: grid of 2^31 points or more");
    poisson_free(c);
    vfs_poisson *ps = new vfs_poisson();
    c->ps = ps;
    hipStream_t st = c->stream;
    int rc = 0;
    // how many levels fit: every coarse cell count must halve exactly
    int nl = 1;
    {
        int cx = c->g.mx - 2, cy = c->g.my - 2, cz = c->g.mz - 2;
        while (nl < levels && nl < 8 && cx % 2 == 0 && cy % 2 == 0 && cz % 2 == 0 && cx >= 4 && cy >= 4 && cz >= 4) {
            cx /= 2; cy /= 2; cz /= 2;
            nl++;
        }
    }
    ps->nl = nl;
    ps->L[0].g = c->g;
    ps->L[0].finest = true;
    rc = level_alloc_work(ps, ps->L[0]);
    if (!rc) rc = palloc(ps, (void **)&ps->L[0].xs32, 2 * (size_t)ps->L[0].z.cplane * 4);
    if (!rc) rc = palloc(ps, (void **)&ps->L[0].bs32, 2 * (size_t)ps->L[0].z.cplane * 4);
    if (rc) return rc;
    double *const m0[17] = {c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_ICSI], c->f[VFS_IETA],
                            c->f[VFS_IZET], c->f[VFS_IAJ], c->f[VFS_JCSI], c->f[VFS_JETA], c->f[VFS_JZET], c->f[VFS_JAJ],
                            c->f[VFS_KCSI], c->f[VFS_KETA], c->f[VFS_KZET], c->f[VFS_KAJ], c->f[VFS_CENT]};
    rc = level_assemble(c, ps->L[0], m0, c->f[VFS_NVERT]);
    if (rc) return rc;
    {
        // VFS_POISSON_MF=1 (no blanked point): the finest FP64 operator runs
        // matrix-free from nine face products per point.  Same bits as the
        // stored coefficients, and at 256x192x256 the same time with an FP32
        // x (395 against 400 us) and slower with an FP64 one (506 against
        // 479 us): the stored apply streams its 152 B of coefficients at about
        // 6.3 TB/s, while the matrix-free one is bound by its 37 loads per point
        // (18 face products of four points, 19 neighbours), not by bytes.  The
        // stored operator is kept by default (the smoother set-up and the AMG
        // read it anyway).
        bool nb = false;
        rc = nvert_all_zero(c, &nb);
        if (rc) return rc;
        const char *e = std::getenv("VFS_POISSON_MF");
        if (nb && e && e[0] == '1') {
            PLevel &L = ps->L[0];
            rc = palloc(ps, (void **)&L.fg, 9 * (size_t)L.g.n * sizeof(double));
            if (rc) return rc;
            FaceSet F = {m0[4], m0[5], m0[6], m0[7], m0[8], m0[9], m0[10], m0[11], m0[12], m0[13], m0[14], m0[15]};
            hipLaunchKernelGGL(k_face_g, dim3(blocks_for(L.g.n, PTPB)), dim3(PTPB), 0, c->stream, L.g, F, L.fg);
            VFS_LAUNCH_CHECK("face products");
        }
    }
    for (int l = 1; l < nl; l++) {
        PLevel &Lf = ps->L[l - 1], &Lc = ps->L[l];
        vfs_params pc = c->p;
        pc.mx = (Lf.g.mx - 1 + 1) / 2 + 1;
        pc.my = (Lf.g.my - 1 + 1) / 2 + 1;
        pc.mz = (Lf.g.mz - 1 + 1) / 2 + 1;
        Lc.g = make_grid(pc);
        size_t n = (size_t)Lc.g.n;
        rc = palloc(ps, (void **)&Lc.coor, 3 * n * 8);
        for (int q = 0; q < 17 && !rc; q++) {
            bool vec = !(q == 3 || q == 7 || q == 11 || q == 15);
            rc = palloc(ps, (void **)&Lc.met[q], (vec ? 3 : 1) * n * 8);
        }
        if (!rc) rc = palloc(ps, (void **)&Lc.nvert, n * 8);
        if (!rc) rc = level_alloc_work(ps, Lc);
        if (rc) return rc;
        const double *fcoor = l == 1 ? c->f[VFS_COOR] : Lf.coor;
        const double *fnv = l == 1 ? c->f[VFS_NVERT] : Lf.nvert;
        hipLaunchKernelGGL(k_coarse_coor, dim3(blocks_for(Lc.g.n, PTPB)), dim3(PTPB), 0, st, Lc.g, Lf.g, fcoor, Lc.coor);
        hipLaunchKernelGGL(k_coarse_nvert, dim3(blocks_for(Lc.g.n, PTPB)), dim3(PTPB), 0, st, Lc.g, Lf.g, fnv, Lc.nvert);
        VFS_LAUNCH_CHECK("coarse grid");
        rc = launch_metrics(c, Lc.g, Lc.coor, Lc.met);
        if (!rc) rc = level_assemble(c, Lc, Lc.met, Lc.nvert);
        if (rc) return rc;
    }
    if (nl > 1) rc = level_dense_inverse(c, ps, ps->L[nl - 1]);
    if (!rc) ps->req_levels = levels;
    return rc;
}

static int ensure_krylov(vfs_ctx *c, int restart)
{
    vfs_poisson *ps = c->ps;
    if (ps->restart >= restart) return VFS_OK;
    size_t n = (size_t)c->g.n;
    int rc = palloc(ps, (void **)&ps->V, (size_t)(restart + 1) * n * 8);
    if (!rc) rc = palloc(ps, (void **)&ps->Z, (size_t)restart * n * 8);
    if (!rc) rc = palloc(ps, (void **)&ps->w, n * 8);
    if (!rc) rc = palloc(ps, (void **)&ps->dev_scalars, 4 * MAXV * 8);
    if (!rc) ps->restart = restart;
    return rc;
}

// ---- V-cycle ---------------------------------------------------------------
template <typename T>
struct LV;
template <>
struct LV<float> {
    static float *x(PLevel &L) { return L.x32; }
    static float *b(PLevel &L) { return L.b32; }
    static float *r(PLevel &L) { return L.r32; }
    static const float *coef(PLevel &L) { return L.c32; }
    static const float *sf(PLevel &L) { return L.sc32; }
    static float *rz(PLevel &L) { return L.rz32; }
    static const float *inv(PLevel &L) { return L.inv32; }
};
template <>
struct LV<double> {
    static double *x(PLevel &L) { return L.x64; }
    static double *b(PLevel &L) { return L.b64; }
    static double *r(PLevel &L) { return L.r64; }
    static const double *coef(PLevel &L) { return L.c64; }
    static const double *sf(PLevel &L) { return L.sc64 + SM_LO * 2 * L.z.cplane; }
    static double *rz(PLevel &L) { return L.rz64; }
    static const double *inv(PLevel &L) { return L.inv64; }
};

constexpr int LINE_PREFETCH = 16;

// Which line-relaxation form a level takes.
struct LinePlan {
    bool short_seg, scan, c4, fused;
    bool split;  // the finest level on the one-pass path: x and b in the S layout
    int nseg, sl;
    int lines;   // lines per block of the one-pass kernel (32 or 64)
};
template <typename T>
static LinePlan line_plan(const PLevel &L)
{
    constexpr int SLM = line_slm<T>();
    LinePlan p;
    // VFS_LINE_PATH forces one line-solve form (tests reach every form on
    // small grids): "short" the 8-point FP32 scan, "long" the SLM-point scan,
    // "solve" one thread per line; otherwise the line length chooses
    const char *lp = std::getenv("VFS_LINE_PATH");
    const int force = !lp ? 0 : lp[0] == 's' && lp[1] == 'h' ? 1 : lp[0] == 'l' ? 2 : lp[0] == 's' ? 3 : 0;
    p.short_seg = sizeof(T) == 4 && L.z.njI <= LINE_SLM_SHORT * LINE_MAXSEG;
    const char *fz = std::getenv("VFS_ZEBRA_FUSED");
    const bool fusable = sizeof(T) == 4 && L.compact && zebra_fused_ok(L.g) && !(fz && fz[0] == '0');
    // on the finest level the one-pass kernel takes 16-point segments and 64
    // lines per block up to 256-point lines (a segment row is a whole
    // wavefront; the coarse levels, which fill fewer blocks, keep 8 and 32)
    if (fusable && L.finest && L.z.njI <= line_slm<float>() * 16) p.short_seg = false;
    if (force == 1) p.short_seg = sizeof(T) == 4;
    if (force == 2) p.short_seg = false;
    const int slm = p.short_seg ? LINE_SLM_SHORT : SLM;
    p.nseg = (L.z.njI + slm - 1) / slm;
    p.scan = p.nseg <= LINE_MAXSEG && force != 3;  // lines up to 32 * slm points
    p.sl = (L.z.njI + p.nseg - 1) / p.nseg;
    p.c4 = sizeof(T) == 4 && L.compact;
    // the compact form's right-hand sides are formed inside the line solve
    // (k_zebra_line_c4) unless VFS_ZEBRA_FUSED=0 asks for the two-pass form
    p.fused = fusable && p.scan;
    p.lines = p.fused && !p.short_seg && p.nseg <= 16 ? 64 : 32;
    // VFS_MG_SPLIT=0 keeps the finest level's vectors in the natural layout
    // (as does the stored-residual restriction, VFS_MG_FUSED_RESTRICT=0)
    const char *fs = std::getenv("VFS_MG_SPLIT"), *fr = std::getenv("VFS_MG_FUSED_RESTRICT");
    // (and a finest level that is also a dense coarsest level keeps it)
    p.split = p.fused && L.finest && L.xs32 && L.nr == 0 && !(fs && fs[0] == '0') && !(fr && fr[0] == '0');
    return p;
}

// sweeps of zebra line Gauss-Seidel on level L.  x_zero: the iterate is zero
// on entry and need not be read by the first half-sweep (one-pass form only;
// the caller then skips the zero fill of the rows, the ghosts being zero).
template <typename T>
static void smooth(vfs_ctx *c, PLevel &L, int sweeps, bool x_zero = false)
{
    hipStream_t st = c->stream;
    const long long lines = (long long)L.z.nkI * L.z.nh;
    const bool fine = L.g.n == c->g.n;
    constexpr int SLM = line_slm<T>();
    const LinePlan P = line_plan<T>(L);
    const bool short_seg = P.short_seg, scan = P.scan, c4 = P.c4, fused = P.fused;
    const int nseg = P.nseg, sl = P.sl;
    const h4_t *sh4 = reinterpret_cast<const h4_t *>(L.sh16);
    const unsigned scan_blocks = (unsigned)((lines + LINE_LINES - 1) / LINE_LINES);
    for (int s = 0; s < sweeps; s++)
        for (int col = 0; col < 2; col++) {
            if (fused) {
                const bool xz = x_zero && s == 0 && col == 0;
                if (fine) prof_begin(c, K_LRELAXC32);
                constexpr int SL = LINE_SLM_SHORT, LG = line_slm<float>();
                const bool w64 = P.lines == 64;  // (long segments only)
                auto k = P.split
                             ? (short_seg ? (xz ? k_zebra_line_c4<SL, true, true, 32> : k_zebra_line_c4<SL, false, true, 32>)
                                : w64     ? (xz ? k_zebra_line_c4<LG, true, true, 64> : k_zebra_line_c4<LG, false, true, 64>)
                                          : (xz ? k_zebra_line_c4<LG, true, true, 32> : k_zebra_line_c4<LG, false, true, 32>))
                             : (short_seg ? (xz ? k_zebra_line_c4<SL, true, false, 32> : k_zebra_line_c4<SL, false, false, 32>)
                                : w64     ? (xz ? k_zebra_line_c4<LG, true, false, 64> : k_zebra_line_c4<LG, false, false, 64>)
                                          : (xz ? k_zebra_line_c4<LG, true, false, 32> : k_zebra_line_c4<LG, false, false, 32>));
                const unsigned blocks = (unsigned)((lines + P.lines - 1) / P.lines);
                hipLaunchKernelGGL(k, dim3(blocks), dim3(P.lines * nseg), 0, st, L.g, L.z, col, nseg, sl, sh4, L.sd32,
                                   P.split ? L.bs32 : L.b32, L.sc32, P.split ? L.xs32 : L.x32);
                if (fine) prof_end(c);
                continue;
            }
            if (fine) prof_begin(c, sizeof(T) == 4 ? (c4 ? K_LINEC32 : K_LINE32) : K_LINE64);
            if constexpr (sizeof(T) == 4) {
                if (c4)
                    hipLaunchKernelGGL(k_zebra_rhs_c4, dim3(blocks_for(L.z.cplane, PTPB)), dim3(PTPB), 0, st, L.g, L.z, col, sh4,
                                       L.sd32, L.x32, L.b32, L.rz32);
                else
                    hipLaunchKernelGGL(k_zebra_rhs_h2,
                                       dim3(blocks_for((long long)L.z.nkI * ((L.z.njI + 1) / 2) * L.z.nh, PTPB)), dim3(PTPB),
                                       0, st, L.g, L.z, col, L.sh16, L.sd32, L.x32, L.b32, L.rz32);
            } else
                hipLaunchKernelGGL(k_zebra_rhs<double>, dim3(blocks_for(L.z.cplane, PTPB)), dim3(PTPB), 0, st, L.g, L.z,
                                   col, L.sc64, L.x64, L.b64, L.rz64);
            if (fine) prof_end(c);
            if (fine) prof_begin(c, sizeof(T) == 4 ? K_LSOLVE32 : K_LSOLVE64);
            if (scan && short_seg) {
                hipLaunchKernelGGL((k_zebra_scan<T, LINE_SLM_SHORT>), dim3(scan_blocks), dim3(LINE_LINES * nseg), 0, st, L.g, L.z,
                                   col, nseg, sl, LV<T>::sf(L), (const T *)LV<T>::rz(L), LV<T>::x(L));
            } else if (scan) {
                hipLaunchKernelGGL((k_zebra_scan<T, SLM>), dim3(scan_blocks), dim3(LINE_LINES * nseg), 0, st, L.g, L.z, col, nseg,
                                   sl, LV<T>::sf(L), (const T *)LV<T>::rz(L), LV<T>::x(L));
            } else {
                hipLaunchKernelGGL((k_zebra_solve<T, LINE_PREFETCH>), dim3(blocks_for(lines, PTPB)), dim3(PTPB), 0, st, L.g,
                                   L.z, col, LV<T>::sf(L), LV<T>::rz(L), LV<T>::x(L));
            }
            if (fine) prof_end(c);
        }
}

template <typename T>
static void vcycle(vfs_ctx *c, vfs_poisson *ps, int l, const Sweeps &S);

// A coarse level's iterate starts each V-cycle at zero.  When its first
// smoothing half-sweep is the one-pass compact form it reads no x (x_zero)
// and rewrites every interior point, and the coarsest dense solve rewrites
// every row; the other points (ghosts, blanked points) are zero since
// allocation and no kernel writes them, so the zero fill is then not needed.
// Levels on other paths keep it.
template <typename T>
static bool coarse_zero_free(const vfs_poisson *ps, int l, const Sweeps &S)
{
    if (l < 1) return false;
    const PLevel &L = ps->L[l];
    const bool last = l == ps->nl - 1;
    if (last && L.nr > 0) return true;  // dense coarsest level
    return (last ? S.coarse : S.pre_at(l)) >= 1 && line_plan<T>(L).fused;
}

// The coarse part of a V-cycle (zero x on level 1, then levels 1..nl-1) is a
// fixed sequence of small kernels on fixed buffers: it is captured once into
// a HIP graph per hierarchy and smoothing setting and replayed, which cuts the
// per-launch gaps of the ~4 us coarse-level kernels.  It holds kernel nodes
// only.  VFS_MG_GRAPH=0 launches the same kernels directly.
template <typename T>
static void coarse_subcycle(vfs_ctx *c, vfs_poisson *ps, const Sweeps &S)
{
    hipStream_t st = c->stream;
    PLevel &C = ps->L[1];
    auto direct = [&]() {
        if (!coarse_zero_free<T>(ps, 1, S))
            hipLaunchKernelGGL(k_zero<T>, dim3(blocks_for(C.g.n, PTPB)), dim3(PTPB), 0, st, C.g.n, LV<T>::x(C));
        vcycle<T>(c, ps, 1, S);
    };
    const char *env = std::getenv("VFS_MG_GRAPH");
    if (env && env[0] == '0') { direct(); return; }
    SubGraph &sg = ps->sub[sizeof(T) == 4 ? 0 : 1];
    if (!sg.exec || !(sg.key == S)) {
        if (sg.exec) { hipGraphExecDestroy(sg.exec); sg.exec = nullptr; }
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) { direct(); return; }
        direct();
        if (hipStreamEndCapture(st, &graph) != hipSuccess || !graph) { direct(); return; }
        hipError_t e = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) { sg.exec = nullptr; direct(); return; }
        sg.key = S;
    }
    // a replay that cannot be enqueued runs the same kernels directly
    if (hipGraphLaunch(sg.exec, st) != hipSuccess) {
        (void)hipGetLastError();
        direct();
    }
}

template <typename T>
static void vcycle(vfs_ctx *c, vfs_poisson *ps, int l, const Sweeps &S)
{
    hipStream_t st = c->stream;
    PLevel &L = ps->L[l];
    const int pre = S.pre_at(l), post = S.post_at(l);
    const unsigned nb = blocks_for(L.g.n, PTPB);
    // the finest level in the S layout (FP32 one-pass path): its iterate starts
    // unfilled; the first half-sweep takes it as zero (a level with no sweep
    // before its iterate is read gets a zero fill)
    const bool spl = sizeof(T) == 4 && line_plan<T>(L).split;
    const bool last = l == ps->nl - 1;
    if (spl && (last ? S.coarse : pre) < 1)
        hipLaunchKernelGGL(k_zero<float>, dim3(blocks_for(2 * L.z.cplane, PTPB)), dim3(PTPB), 0, st, 2 * L.z.cplane, L.xs32);
    if (last) {
        if (L.nr > 0)
            hipLaunchKernelGGL(k_dense_solve<T>, dim3(blocks_for(L.nr, DENSE_ROWS)), dim3(DENSE_ROWS * DENSE_SPLIT), 0, st, L.nr, L.ridx,
                               LV<T>::inv(L), LV<T>::b(L), LV<T>::x(L));
        else
            smooth<T>(c, L, S.coarse, spl || coarse_zero_free<T>(ps, l, S));
        return;
    }
    smooth<T>(c, L, pre, spl || coarse_zero_free<T>(ps, l, S));
    // a compact level's V-cycle works on the axis part of its operator: the
    // residual it restricts takes the centre and the six axis neighbours.
    // The residual is formed inside the restriction (k_residual_restrict)
    // unless VFS_MG_FUSED_RESTRICT=0 asks for the stored-residual form.
    PLevel &C = ps->L[l + 1];
    const unsigned nbc = blocks_for(C.g.n, PTPB);
    const char *fr = std::getenv("VFS_MG_FUSED_RESTRICT");
    const bool axis = sizeof(T) == 4 && L.compact;
    if (spl) {
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((k_residual_restrict<float, 7, true>), dim3(nbc), dim3(PTPB), 0, st, C.g, L.g, L.c32, C.row, L.row,
                               (const float *)L.xs32, (const float *)L.bs32, C.b32, L.z);
    } else if (!(fr && fr[0] == '0')) {
        if (axis)
            hipLaunchKernelGGL((k_residual_restrict<T, 7>), dim3(nbc), dim3(PTPB), 0, st, C.g, L.g, LV<T>::coef(L), C.row, L.row,
                               (const T *)LV<T>::x(L), (const T *)LV<T>::b(L), LV<T>::b(C), L.z);
        else
            hipLaunchKernelGGL((k_residual_restrict<T, NST>), dim3(nbc), dim3(PTPB), 0, st, C.g, L.g, LV<T>::coef(L), C.row,
                               L.row, (const T *)LV<T>::x(L), (const T *)LV<T>::b(L), LV<T>::b(C), L.z);
    } else {
        if (axis)
            hipLaunchKernelGGL((k_apply<T, T, 7>), dim3(nb), dim3(PTPB), 0, st, L.g, LV<T>::coef(L), L.row, LV<T>::x(L),
                               LV<T>::b(L), LV<T>::r(L));
        else
            hipLaunchKernelGGL(k_apply<T>, dim3(nb), dim3(PTPB), 0, st, L.g, LV<T>::coef(L), L.row, LV<T>::x(L), LV<T>::b(L),
                               LV<T>::r(L));
        hipLaunchKernelGGL(k_restrict<T>, dim3(nbc), dim3(PTPB), 0, st, C.g, L.g, C.row, L.row, LV<T>::r(L), LV<T>::b(C));
    }
    if (l == 0) {
        coarse_subcycle<T>(c, ps, S);
    } else {
        if (!coarse_zero_free<T>(ps, l + 1, S))
            hipLaunchKernelGGL(k_zero<T>, dim3(blocks_for(C.g.n, PTPB)), dim3(PTPB), 0, st, C.g.n, LV<T>::x(C));
        vcycle<T>(c, ps, l + 1, S);
    }
    if (spl) {
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL(k_prolong_split<float>, dim3(blocks_for(2 * L.z.cplane, PTPB)), dim3(PTPB), 0, st, C.g, L.g, L.z,
                               L.row, (const float *)C.x32, L.xs32);
    } else {
        hipLaunchKernelGGL(k_prolong<T>, dim3(nb), dim3(PTPB), 0, st, C.g, L.g, L.row, LV<T>::x(C), LV<T>::x(L));
    }
    smooth<T>(c, L, post);
}

// z = M^{-1} v (V-cycle from zero); FP32 or FP64 hierarchy
// With z32 the FP32 V-cycle's finest iterate is z32 itself (no FP64 copy);
// the FP64 V-cycle iterates in z directly.  The finest level's kernels take
// the iterate pointer at launch (the captured coarse levels never see it).
// y = A x or r = b - A x with the finest-level FP64 operator: matrix-free
// when the face products are there, else from the stored coefficients
template <typename X>
static void apply0(vfs_ctx *c, const X *x, const double *b, double *y)
{
    PLevel &L = c->ps->L[0];
    const unsigned nb = blocks_for(L.g.n, PTPB);
    if (L.fg) {
        const FaceG F{L.fg, L.g.n};
        hipLaunchKernelGGL((k_apply_mf<double, X>), dim3(nb), dim3(PTPB), 0, c->stream, L.g, F, c->p.poisson_threshold,
                           L.row, x, b, y);
    } else {
        hipLaunchKernelGGL((k_apply<double, X>), dim3(nb), dim3(PTPB), 0, c->stream, L.g, L.c64, L.row, x, b, y);
    }
}

// neighbour point of point p in stencil slot s on the level-0 grid
static long long amg_nbr(const void *grid, long long p, int s)
{
    const Grid &g = *static_cast<const Grid *>(grid);
    const int i = (int)(p % g.mx), j = (int)((p / g.mx) % g.my), k = (int)(p / g.plane);
    return nbr_idx(g, k, j, i, s);
}

// The AMG hierarchy of the level-0 operator for the requested threshold
// (built on the host from the assembled FP64 coefficients).
static int amg_ensure(vfs_ctx *c, const vfs_solve_opts *o)
{
    vfs_poisson *ps = c->ps;
    const double theta = o->amg_thresh > 0 && o->amg_thresh < 1 ? o->amg_thresh : 0.75;
    if (ps->amg && amg_theta(ps->amg) == theta) return VFS_OK;
    amg_free(ps->amg);
    ps->amg = nullptr;
    PLevel &L = ps->L[0];
    const long long n = L.g.n;
    std::vector<double> c64((size_t)NST * n);
    std::vector<unsigned char> row(n);
    VFS_HIP(hipStreamSynchronize(c->stream));
    VFS_HIP(hipMemcpy(c64.data(), L.c64, c64.size() * sizeof(double), hipMemcpyDeviceToHost));
    VFS_HIP(hipMemcpy(row.data(), L.row, n, hipMemcpyDeviceToHost));
    return amg_build(c64.data(), row.data(), n, amg_nbr, &L.g, theta, &ps->amg);
}

static Sweeps sweeps_of(const vfs_solve_opts *o)
{
    return Sweeps{o->pre, o->post, o->coarse_sweeps, o->deep_level, o->deep_pre, o->deep_post};
}

static int precondition(vfs_ctx *c, const vfs_solve_opts *o, const double *v, double *z, float *z32 = nullptr,
                        bool b32_ready = false)
{
    vfs_poisson *ps = c->ps;
    hipStream_t st = c->stream;
    PLevel &L = ps->L[0];
    const long long n = L.g.n;
    const unsigned nb = blocks_for(n, PTPB);
    if (o->amg) return amg_apply(st, ps->amg, n, v, z);
    if (o->mg_fp32) {
        // on the one-pass path the V-cycle's finest vectors are in the S
        // layout: b converted into it, the iterate converted back into the
        // natural x (zero at the ghosts) at the end
        const bool spl = line_plan<float>(L).split;
        if (!b32_ready) {
            if (spl) hipLaunchKernelGGL(k_convert_split, dim3(nb), dim3(PTPB), 0, st, L.g, L.z, v, L.bs32);
            else hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(PTPB), 0, st, n, v, L.b32);
        }
        float *x0 = L.x32;
        if (z32) L.x32 = z32;
        if (!spl) hipLaunchKernelGGL(k_zero<float>, dim3(nb), dim3(PTPB), 0, st, n, L.x32);
        vcycle<float>(c, ps, 0, sweeps_of(o));
        if (spl) hipLaunchKernelGGL(k_split_to_natural, dim3(nb), dim3(PTPB), 0, st, L.g, L.z, (const float *)L.xs32, L.x32);
        if (!z32) hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(PTPB), 0, st, n, L.x32, z);
        L.x32 = x0;
    } else {
        hipMemcpyAsync(L.b64, v, n * 8, hipMemcpyDeviceToDevice, st);
        double *x0 = L.x64;
        L.x64 = z;
        hipLaunchKernelGGL(k_zero<double>, dim3(nb), dim3(PTPB), 0, st, n, L.x64);
        vcycle<double>(c, ps, 0, sweeps_of(o));
        L.x64 = x0;
    }
    VFS_LAUNCH_CHECK("v-cycle");
    return VFS_OK;
}

// dots of w with V_0..V_{nv-1} plus <w,w>, returned on the host
static int multidot(vfs_ctx *c, int nv, const double *V, const double *w, double *out)
{
    hipStream_t st = c->stream;
    const long long n = c->g.n;
    const dim3 gb(RED_BLOCKS), tb(RED_THREADS);
    if (nv <= 4) hipLaunchKernelGGL((k_multidot<4, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 8) hipLaunchKernelGGL((k_multidot<8, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 16) hipLaunchKernelGGL((k_multidot<16, 2>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else hipLaunchKernelGGL((k_multidot<MAXV - 1, 1>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    hipLaunchKernelGGL(k_finish_many, dim3(nv + 1), dim3(RED_THREADS), 0, st, nv, c->red, c->red + MAXV * RED_BLOCKS);
    VFS_LAUNCH_CHECK("multidot");
    VFS_HIP(hipMemcpyAsync(c->red_host, c->red + MAXV * RED_BLOCKS, (nv + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    std::memcpy(out, c->red_host, (nv + 1) * sizeof(double));
    return VFS_OK;
}

// w -= sum_v h_v V_v (h in dev_scalars) with the partials of k_multiaxpy
template <bool DOTS>
static void launch_multiaxpy(vfs_ctx *c, int nv, const double *V, double *w, const double *h = nullptr,
                             const double *gww = nullptr, const double *gww0 = nullptr)
{
    hipStream_t st = c->stream;
    const long long n = c->g.n;
    if (!h) h = c->ps->dev_scalars;
    const dim3 gb(RED_BLOCKS), tb(RED_THREADS);
    if (nv <= 4) hipLaunchKernelGGL((k_multiaxpy<4, 4, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, gww, gww0);
    else if (nv <= 8) hipLaunchKernelGGL((k_multiaxpy<8, 4, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, gww, gww0);
    else if (nv <= 16) hipLaunchKernelGGL((k_multiaxpy<16, 2, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, gww, gww0);
    else hipLaunchKernelGGL((k_multiaxpy<MAXV - 1, 1, DOTS>), gb, tb, 0, st, n, nv, h, V, n, w, c->red, gww, gww0);
}

// One classical Gram-Schmidt step of w against V_0..V_{nv-1} with the
// re-orthogonalisation pass when <w, w> drops below a quarter, and
// V_nv = w / ||w||, all on the device: the coefficients of each pass are the
// previous pass's device results, so the host synchronises once.  out (host)
// receives h (nv), ww0, the second-pass dots (nv), ww1 and the final <w, w>
// at [0..nv], [40..40+nv], [80].
constexpr int GS_D2 = 40, GS_D3 = 80;
static_assert(GS_D2 >= MAXV && GS_D3 >= GS_D2 + MAXV, "Gram-Schmidt result slots");
static int gs_step(vfs_ctx *c, int nv, const double *V, double *w, double *vnext, double *out, float *vnext32,
                   ZGrid zs = {})
{
    hipStream_t st = c->stream;
    const long long n = c->g.n;
    double *R = c->red + MAXV * RED_BLOCKS;
    const dim3 gb(RED_BLOCKS), tb(RED_THREADS);
    if (nv <= 4) hipLaunchKernelGGL((k_multidot<4, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 8) hipLaunchKernelGGL((k_multidot<8, 4>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else if (nv <= 16) hipLaunchKernelGGL((k_multidot<16, 2>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    else hipLaunchKernelGGL((k_multidot<MAXV - 1, 1>), gb, tb, 0, st, n, nv, V, n, w, c->red);
    hipLaunchKernelGGL(k_finish_many, dim3(nv + 1), dim3(RED_THREADS), 0, st, nv, c->red, R);
    launch_multiaxpy<true>(c, nv, V, w, R);
    hipLaunchKernelGGL(k_finish_many, dim3(nv + 1), dim3(RED_THREADS), 0, st, nv, c->red, R + GS_D2);
    launch_multiaxpy<false>(c, nv, V, w, R + GS_D2, R + GS_D2 + nv, R + nv);
    hipLaunchKernelGGL(k_finish_gated, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, R + GS_D2 + nv, R + nv,
                       R + GS_D3);
    hipLaunchKernelGGL(k_scale_rnorm, dim3(blocks_for(n, PTPB)), dim3(PTPB), 0, st, n, R + GS_D3, w, vnext, vnext32, c->g, zs);
    VFS_LAUNCH_CHECK("gram-schmidt");
    VFS_HIP(hipMemcpyAsync(c->red_host, R, (GS_D3 + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    VFS_HIP(hipStreamSynchronize(st));
    std::memcpy(out, c->red_host, (GS_D3 + 1) * sizeof(double));
    return VFS_OK;
}

static int remove_mean(vfs_ctx *c, const unsigned char *row, double *x)
{
    hipStream_t st = c->stream;
    const long long n = c->g.n;
    hipLaunchKernelGGL(k_row_sum, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, row, x, c->red);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + 2 * RED_BLOCKS);
    hipLaunchKernelGGL(finish_reduce<OpSum>, dim3(1), dim3(RED_THREADS), 0, st, c->red + RED_BLOCKS, RED_BLOCKS,
                       c->red + 2 * RED_BLOCKS + 1);
    hipLaunchKernelGGL(k_sub_mean, dim3(blocks_for(n, PTPB)), dim3(PTPB), 0, st, n, row, c->red + 2 * RED_BLOCKS, x);
    VFS_LAUNCH_CHECK("remove mean");
    return VFS_OK;
}

int poisson_apply(vfs_ctx *c, const double *x, double *y)
{
    int rc = poisson_setup(c, c->ps ? c->ps->nl : 4);
    if (rc) return rc;
    apply0(c, x, (const double *)nullptr, y);
    VFS_LAUNCH_CHECK("poisson apply");
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

// FGMRES; with a network: warm start and first preconditioner stage from the
// network, and *fell_back set (solve abandoned) when after three iterations
// the residual is above half of its value after the warm start
static int fgmres(vfs_ctx *c, const vfs_solve_opts *o, vfs_unet *net, vfs_solve_stats *stats, bool *fell_back)
{
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    const int m = o->restart > 0 ? (o->restart < MAXV - 1 ? o->restart : MAXV - 2) : 20;
    int rc = poisson_setup(c, o->levels > 0 ? o->levels : 4);
    if (!rc) rc = ensure_krylov(c, m);
    if (rc) return rc;
    vfs_poisson *ps = c->ps;
    hipStream_t st = c->stream;
    PLevel &L = ps->L[0];
    const long long n = L.g.n;
    const unsigned nb = blocks_for(n, PTPB);
    double *x = c->f[VFS_PHI], *b = c->f[VFS_B];
    double *V = ps->V, *Z = ps->Z, *w = ps->w;
    // the FP32 V-cycle's outputs are kept as FP32 vectors in Z's memory (exact);
    // the network's direction is rounded to FP32 as well, and A is applied to
    // the rounded vector, so the FGMRES relation holds for the stored z exactly
    const bool z32 = o->mg_fp32 != 0 && !o->amg;
    if (o->amg) {
        rc = amg_ensure(c, o);
        if (rc) return rc;
    }
    float *Z32 = reinterpret_cast<float *>(Z);
    if (!o->use_guess) hipLaunchKernelGGL(k_zero<double>, dim3(nb), dim3(PTPB), 0, st, n, x);
    double d[MAXV + 1];
    if (net) {
        // x += alpha y, y = net(b - A x), alpha = <r, A y> / <A y, A y>
        apply0(c, (const double *)x, b, w);
        rc = vfs_unet_forward(net, w, Z);
        if (rc) return rc;
        apply0(c, (const double *)Z, (const double *)nullptr, V + n);
        rc = multidot(c, 1, V + n, w, d);
        if (rc) return rc;
        const double rAy = d[0];
        rc = multidot(c, 0, V, V + n, d);
        if (rc) return rc;
        c->red_host[0] = d[0] > 0 ? rAy / d[0] : 0.0;
        VFS_HIP(hipMemcpyAsync(ps->dev_scalars, c->red_host, sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_multiupdate<double>, dim3(nb), dim3(PTPB), 0, st, n, 1, ps->dev_scalars, (const double *)Z, n, x);
        VFS_LAUNCH_CHECK("learned warm start");
    }
    double rel_start = -1;
    // ||b||: multidot with no basis vectors returns <b,b>
    rc = multidot(c, 0, V, b, d);
    if (rc) return rc;
    double bnorm = std::sqrt(d[0]);
    if (bnorm == 0) bnorm = 1;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    int it = 0;
    double rel = 1;
    const double tol = o->tol > 0 ? o->tol : 1e-8;
    const int maxit = o->maxit > 0 ? o->maxit : 200;
    bool converged = false;
    const char *tr = std::getenv("VFS_SOLVER_TRACE");
    const bool trace = tr && tr[0] == '1';
    // the network gives the warm start; with VFS_LEARNED_DIRECTION=1 also the
    // first Krylov direction (on a held-out 256x192x256 channel that took one
    // FGMRES iteration more than the V-cycle's direction: 7 against 6)
    const char *ld = std::getenv("VFS_LEARNED_DIRECTION");
    const bool net_dir = net && ld && ld[0] == '1';
    while (it < maxit) {
        // r = b - A x into V_0
        apply0(c, (const double *)x, b, V);
        rc = multidot(c, 0, V, V, d);
        if (rc) return rc;
        double beta = std::sqrt(d[0]);
        rel = beta / bnorm;
        if (rel_start < 0) rel_start = rel;
        if (trace) std::fprintf(stderr, "fgmres restart at it %d: true relres %.3e (bnorm %.3e)\n", it, rel, bnorm);
        if (rel <= tol) { converged = true; break; }
        c->red_host[0] = 1.0 / beta;
        VFS_HIP(hipMemcpyAsync(ps->dev_scalars, c->red_host, sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scale_dev, dim3(nb), dim3(PTPB), 0, st, n, ps->dev_scalars, 0, V, V);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int jj;
        for (jj = 0; jj < m && it < maxit; jj++, it++) {
            double *vj = V + (size_t)jj * n, *zj = Z + (size_t)jj * n;
            float *zj32 = Z32 + (size_t)jj * n;
            if (net_dir && it == 0) {
                rc = vfs_unet_forward(net, vj, z32 ? w : zj);
                if (!rc && z32) hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(PTPB), 0, st, n, (const double *)w, zj32);
            } else {
                rc = precondition(c, o, vj, zj, z32 ? zj32 : nullptr, z32 && jj > 0);
            }
            if (rc) return rc;
            prof_begin(c, K_APPLY64);
            if (z32)
                apply0(c, (const float *)zj32, (const double *)nullptr, w);
            else
                apply0(c, (const double *)zj, (const double *)nullptr, w);
            prof_end(c);
            // classical Gram-Schmidt, re-orthogonalised when the norm drops;
            // the first pass also forms the dot products of the second
            double gs[GS_D3 + 1];
            // with FP32 V-cycles the new basis vector is also written as the
            // next V-cycle's FP32 right-hand side
            // (in the S layout when the V-cycle's finest level takes it)
            const bool spl = z32 && line_plan<float>(L).split;
            rc = gs_step(c, jj + 1, V, w, V + (size_t)(jj + 1) * n, gs, z32 ? (spl ? L.bs32 : L.b32) : nullptr,
                         spl ? L.z : ZGrid{});
            if (rc) return rc;
            double h[MAXV];
            for (int v = 0; v <= jj; v++) h[v] = gs[v];
            const double ww0 = gs[jj + 1], ww1 = gs[GS_D2 + jj + 1];
            double ww = ww1;
            if (ww1 < 0.25 * ww0) {
                for (int v = 0; v <= jj; v++) h[v] += gs[GS_D2 + v];
                ww = gs[GS_D3];
            }
            double hn = std::sqrt(ww > 0 ? ww : 0);
            for (int v = 0; v <= jj; v++) H[(size_t)v * m + jj] = h[v];
            H[(size_t)(jj + 1) * m + jj] = hn;
            for (int i = 0; i < jj; i++) {
                double tmp = cs[i] * H[(size_t)i * m + jj] + sn[i] * H[(size_t)(i + 1) * m + jj];
                H[(size_t)(i + 1) * m + jj] = -sn[i] * H[(size_t)i * m + jj] + cs[i] * H[(size_t)(i + 1) * m + jj];
                H[(size_t)i * m + jj] = tmp;
            }
            double a = H[(size_t)jj * m + jj], bb = H[(size_t)(jj + 1) * m + jj];
            double den = std::sqrt(a * a + bb * bb);
            cs[jj] = a / den; sn[jj] = bb / den;
            H[(size_t)jj * m + jj] = den;
            H[(size_t)(jj + 1) * m + jj] = 0;
            gv[jj + 1] = -sn[jj] * gv[jj];
            gv[jj] = cs[jj] * gv[jj];
            if (trace) std::fprintf(stderr, "  it %d: est relres %.3e  h_{j+1,j} %.3e  ww0 %.3e ww %.3e\n", it,
                                    std::fabs(gv[jj + 1]) / bnorm, hn, ww0, ww);
            if (std::fabs(gv[jj + 1]) / bnorm <= tol) { jj++; it++; break; }
            // (written so that a NaN residual, from a network that returns non-finite values, also falls back)
            if (net && it == 2 && !(std::fabs(gv[jj + 1]) / bnorm <= 0.5 * rel_start)) {
                *fell_back = true;
                return VFS_OK;
            }
        }
        for (int i = jj - 1; i >= 0; i--) {
            double s = gv[i];
            for (int q = i + 1; q < jj; q++) s -= H[(size_t)i * m + q] * y[q];
            y[i] = s / H[(size_t)i * m + i];
        }
        std::memcpy(c->red_host, y.data(), jj * sizeof(double));
        VFS_HIP(hipMemcpyAsync(ps->dev_scalars, c->red_host, jj * sizeof(double), hipMemcpyHostToDevice, st));
        if (z32)
            hipLaunchKernelGGL(k_multiupdate<float>, dim3(nb), dim3(PTPB), 0, st, n, jj, ps->dev_scalars, (const float *)Z32, n, x);
        else
            hipLaunchKernelGGL(k_multiupdate<double>, dim3(nb), dim3(PTPB), 0, st, n, jj, ps->dev_scalars, (const double *)Z, n, x);
        VFS_LAUNCH_CHECK("fgmres update");
    }
    if (!converged) {
        apply0(c, (const double *)x, b, w);
        rc = multidot(c, 0, V, w, d);
        if (rc) return rc;
        rel = std::sqrt(d[0]) / bnorm;
        converged = rel <= tol;
    }
    rc = remove_mean(c, L.row, x);  // poisson_hypre.c:608
    if (rc) return rc;
    VFS_HIP(hipStreamSynchronize(st));
    if (stats) {
        stats->iterations = it;
        stats->relres = rel;
        stats->ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        stats->fallback = 0;
    }
    if (!converged) return set_error(VFS_ENOCONV, "Poisson FGMRES did not reach the tolerance");
    return VFS_OK;
}

int poisson_solve(vfs_ctx *c, const vfs_solve_opts *o, vfs_solve_stats *stats)
{
    bool fb = false;
    return fgmres(c, o, nullptr, stats, &fb);
}

int poisson_solve_learned(vfs_ctx *c, vfs_unet *net, const vfs_solve_opts *o, vfs_solve_stats *stats)
{
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    const int m = o->restart > 0 ? (o->restart < MAXV - 1 ? o->restart : MAXV - 2) : 20;
    if (m < 4) return set_error(VFS_EINVAL, "learned Poisson solve: restart must be at least 4");
    int rc = poisson_setup(c, o->levels > 0 ? o->levels : 4);
    if (!rc) rc = ensure_krylov(c, m);
    if (rc) return rc;
    const long long n = c->g.n;
    // keep the initial guess for a fallback (in the last Krylov direction slot)
    double *saved = c->ps->Z + (size_t)(m - 1) * n;
    VFS_HIP(hipMemcpyAsync(saved, c->f[VFS_PHI], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    bool fb = false;
    vfs_solve_stats s1;
    std::memset(&s1, 0, sizeof(s1));
    rc = fgmres(c, o, net, &s1, &fb);
    if (rc && rc != VFS_ENOCONV) return rc;
    if (!fb) {
        if (stats) { *stats = s1; stats->ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }
        return rc;
    }
    VFS_HIP(hipMemcpyAsync(c->f[VFS_PHI], saved, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    vfs_solve_opts o2 = *o;
    vfs_solve_stats s2;
    std::memset(&s2, 0, sizeof(s2));
    rc = fgmres(c, &o2, nullptr, &s2, &fb);
    if (stats) {
        *stats = s2;
        stats->iterations += 3;
        stats->fallback = 1;
        stats->ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    return rc;
}

}  // namespace vfs
